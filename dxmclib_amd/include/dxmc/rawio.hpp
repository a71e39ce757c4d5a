// Raw binary arrays: phantoms and dose maps kept as the reference's
// validation program keeps them (validation.cpp:128-135, 1280-1306). A file
// holds exactly the array's bytes in host byte order, with no header; a
// reader is told the element count and refuses a file of any other size.
#pragma once

#include <cstddef>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

namespace dxmc {

// the array in 'path', or nullptr if the file is missing, empty or not
// exactly n elements long
template <typename T>
std::shared_ptr<std::vector<T>> readBinaryArray(const std::string& path, std::size_t n)
{
    std::unique_ptr<std::FILE, int (*)(std::FILE*)> f(std::fopen(path.c_str(), "rb"), &std::fclose);
    if (!f || n == 0)
        return nullptr;
    if (std::fseek(f.get(), 0, SEEK_END) != 0)
        return nullptr;
    const long size = std::ftell(f.get());
    if (size < 0 || static_cast<unsigned long>(size) != n * sizeof(T) || std::fseek(f.get(), 0, SEEK_SET) != 0)
        return nullptr;
    auto data = std::make_shared<std::vector<T>>(n);
    if (std::fread(data->data(), sizeof(T), n, f.get()) != n)
        return nullptr;
    return data;
}

// writes the array's bytes to 'path'; false if the file cannot be written
template <typename T>
bool saveBinaryArray(const std::vector<T>& data, const std::string& path)
{
    std::unique_ptr<std::FILE, int (*)(std::FILE*)> f(std::fopen(path.c_str(), "wb"), &std::fclose);
    if (!f)
        return false;
    if (!data.empty() && std::fwrite(data.data(), sizeof(T), data.size(), f.get()) != data.size())
        return false;
    return std::fflush(f.get()) == 0;
}

}
