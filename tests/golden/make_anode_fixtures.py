"""Write tests/golden/anode_reference.npz: the numeric data of the reference's
tube spectrum model and its published validation spectra, read as text from
a reference checkout (default /root/reference; DXMC_REFERENCE to override):

  * electron-penetration tables, Thomson-Whiddington constants and tungsten
    K lines (include/dxmc/betheHeitlerCrossSection.hpp:36-171, 394-407);
  * TG-195 RQR-8 (100 kVp) and RQR-9 (120 kVp) spectra, W anode 11 deg with
    2.708 / 2.861 mm Al (validation/validation.cpp:137-170).

tests/test_tube_spectrum.py checks the product's spectrum against the
oracle's direct integration over these tables, and against the published
spectra.

    python tests/golden/make_anode_fixtures.py
"""
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE = os.environ.get("DXMC_REFERENCE", "/root/reference")


def _numbers(text):
    return np.array([float(v) for v in text.replace("\n", " ").split(",") if v.strip()])


def _array(src, name):
    m = re.search(r"\b" + name + r"\(\)\s*\{\s*std::array<T,[^=]*?>\s*res\s*=\s*\{([^}]*)\}", src, re.S)
    if not m:
        raise SystemExit(f"{name} not found")
    return _numbers(m.group(1))


def main():
    bh = open(os.path.join(REFERENCE, "include", "dxmc", "betheHeitlerCrossSection.hpp")).read()
    out = {
        "u": _array(bh, "CP100_u"),
        "depth_f": _array(bh, "CP100_F_x"),
        "depth_m": _array(bh, "CP100_M_x"),
        "tw_voltage": _array(bh, "THOMSONWIDDINGTONCONSTANT_T"),
        "tw_constant": _array(bh, "THOMSONWIDDINGTONCONSTANT_C"),
    }
    nu, nx = len(out["u"]), len(out["depth_f"])
    out["density_f"] = _array(bh, "CP100_F").reshape(nu, nx)
    out["density_m"] = _array(bh, "CP100_M").reshape(nu, len(out["depth_m"]))
    m = re.search(r"k_edge_energies\s*\{([^}]*)\}", bh)
    out["k_energy"] = _numbers(m.group(1))
    m = re.search(r"k_edge_fractions\s*\{([^}]*)\}", bh)
    out["k_fraction"] = _numbers(m.group(1))
    val = open(os.path.join(REFERENCE, "validation", "validation.cpp")).read()
    for kv in (100, 120):
        m = re.search(rf"TG195_{kv}KV_raw\(\{{([^}}]*)\}}", val)
        raw = _numbers(m.group(1))
        out[f"tg195_{kv}_energy"] = raw[0::2]
        out[f"tg195_{kv}_weight"] = raw[1::2]
    np.savez_compressed(os.path.join(HERE, "anode_reference.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
