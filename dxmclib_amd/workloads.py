"""Synthetic worlds and sources of the benchmark configurations.

water_pencil   64^3 x 1 mm water, 60 keV pencil beam along +z into the centre
               of one face, 10 exposures x 1e5 histories, seed 42
dx_thorax      256x256x128 x 1.5 mm: water elliptic cylinder (300x200 mm)
               with a 10 mm adipose ring and 8 seeded bone / lung ellipsoids
               in air; DX 120 kVp, 2.5 mm Al, 300x300 mm field at 1000 mm,
               1000 exposures x 1e5 histories
ct_body        512x512x300 x 0.8 mm seeded ellipsoid body (soft tissue,
               adipose skin, lungs, spine, ribs, water stomach, air), density
               jittered +-2 %; CT spiral 120 kVp, 6 mm Al, bowtie, 40 mm
               collimation, pitch 1, 0.5 deg steps over 200 mm (3600
               exposures), 1e9 histories
serving_stream 256 Poisson requests of water/bone cubes 32-256 voxels, DX or
               CT axial 80-140 kVp, 1e6-1e8 histories
whole_body     1024x1024x600 x 0.7 mm, 8 materials, CT axial 360 x 1 deg,
               1e10 histories

Each world is returned as (material u8 [z,y,x], density f32 [z,y,x],
spacing (x,y,z) mm, material names) and each source as a SourceSpec.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import capi
from .backend import source_spec

WATER = "Water, Liquid"
AIR = "Air, Dry (near sea level)"
BONE = "Bone, Cortical (ICRU)"
LUNG = "Lung (ICRU)"
ADIPOSE = "Adipose Tissue (ICRP)"
SOFT = "Tissue, Soft (ICRU-44)"


@dataclass
class Workload:
    name: str
    material: np.ndarray
    density: np.ndarray
    spacing: tuple
    names: list
    spec: object
    histories: int
    description: str = ""
    extra: dict = field(default_factory=dict)

    @property
    def voxels(self):
        return int(self.material.size)


def _grid(shape, spacing):
    """Voxel-centre coordinates (z, y, x) in mm of a grid centred on the origin."""
    nz, ny, nx = shape
    sx, sy, sz = spacing
    x = (np.arange(nx) - nx / 2 + 0.5) * sx
    y = (np.arange(ny) - ny / 2 + 0.5) * sy
    z = (np.arange(nz) - nz / 2 + 0.5) * sz
    return z, y, x


def water_pencil(histories_per_exposure=100_000, exposures=10):
    mat = np.zeros((64, 64, 64), dtype=np.uint8)
    dens = np.ones((64, 64, 64), dtype=np.float32)
    spec = source_spec(capi.SOURCE_PENCIL, position=(0.0, 0.0, -64.0), direction_cosines=(1, 0, 0, 0, 1, 0), photon_energy=60.0,
                       histories_per_exposure=histories_per_exposure, total_exposures=exposures)
    return Workload("water_pencil_60keV_64cube", mat, dens, (1.0, 1.0, 1.0), [WATER], spec, histories_per_exposure * exposures,
                    "64^3 x 1 mm water, 60 keV pencil beam, 1e6 histories")


def _ellipsoid(z, y, x, c, r):
    return ((x[None, None, :] - c[0]) / r[0]) ** 2 + ((y[None, :, None] - c[1]) / r[1]) ** 2 + ((z[:, None, None] - c[2]) / r[2]) ** 2 <= 1.0


def _paint(mat, z, y, x, c, r, value):
    """mat[_ellipsoid(z, y, x, c, r)] = value, evaluated only over the
    ellipsoid's bounding box (the same voxels, without full-grid temporaries)."""
    sl = []
    for coord, cc, rr in ((z, c[2], r[2]), (y, c[1], r[1]), (x, c[0], r[0])):
        idx = np.nonzero(np.abs(coord - cc) <= rr)[0]
        if idx.size == 0:
            return
        sl.append(slice(int(idx[0]), int(idx[-1]) + 1))
    sub = _ellipsoid(z[sl[0]], y[sl[1]], x[sl[2]], c, r)
    mat[sl[0], sl[1], sl[2]][sub] = value


def dx_thorax(histories_per_exposure=100_000, exposures=1000, seed=7):
    shape = (128, 256, 256)
    sp = (1.5, 1.5, 1.5)
    z, y, x = _grid(shape, sp)
    names = [AIR, WATER, ADIPOSE, BONE, LUNG]
    dens_of = [0.001205, 1.0, 0.95, 1.92, 0.26]
    mat = np.zeros(shape, dtype=np.uint8)
    ell = (x[None, None, :] / 150.0) ** 2 + (y[None, :, None] / 100.0) ** 2
    ell = np.broadcast_to(ell, shape)
    mat[ell <= 1.0] = 2
    inner = (x[None, None, :] / 140.0) ** 2 + (y[None, :, None] / 90.0) ** 2 <= 1.0
    mat[np.broadcast_to(inner, shape)] = 1
    rng = np.random.default_rng(seed)
    for k in range(8):
        while True:
            r = rng.uniform(10, 40, 3)
            c = np.array([rng.uniform(-120, 120), rng.uniform(-70, 70), rng.uniform(-80, 80)])
            # fully inside the water: the ellipsoid's bounding box within the inner ellipse and slab
            corners = [(c[0] + sx * r[0], c[1] + sy * r[1]) for sx in (-1, 1) for sy in (-1, 1)]
            if all((cx / 140.0) ** 2 + (cy / 90.0) ** 2 <= 1.0 for cx, cy in corners) and abs(c[2]) + r[2] < 96:
                break
        _paint(mat, z, y, x, c, r, 3 if k < 4 else 4)
    dens = np.asarray(dens_of, dtype=np.float32)[mat]
    spec = source_spec(capi.SOURCE_DX, tube_voltage=120.0, anode_angle_deg=12.0, energy_resolution=1.0, al_mm=2.5, sdd=1000.0,
                       field_size=(300.0, 300.0), histories_per_exposure=histories_per_exposure, total_exposures=exposures, heel_effect=1)
    return Workload("dx_thorax_120kVp_256x256x128", mat, dens, sp, names, spec, histories_per_exposure * exposures,
                    "DX 120 kVp, 256x256x128 x 1.5 mm water cylinder with bone/lung inserts")


def bowtie_profile():
    a = np.linspace(0.0, 0.45, 64)
    return a, np.exp(-6.0 * a * a)


def ct_spiral_spec(histories_total=1_000_000_000, tube_voltage=120.0, scan_length=200.0, start_z=-100.0):
    exposures = int(scan_length * 2 * math.pi / (40.0 * 1.0 * math.radians(0.5)))
    per = max(1, round(histories_total / exposures))
    ang, w = bowtie_profile()
    spec = source_spec(capi.SOURCE_CT_SPIRAL, position=(0.0, 0.0, start_z), tube_voltage=tube_voltage, anode_angle_deg=12.0, al_mm=6.0,
                       sdd=1190.0, fov=500.0, collimation=40.0, pitch=1.0, exposure_angle_step_deg=0.5, scan_length=scan_length,
                       histories_per_exposure=per, bowtie_angles=ang, bowtie_weights=w, heel_effect=1)
    return spec, exposures, per


def ct_body(histories_total=1_000_000_000, seed=11, shape=(300, 512, 512), spacing=(0.8, 0.8, 0.8)):
    z, y, x = _grid(shape, spacing)
    names = [AIR, SOFT, ADIPOSE, LUNG, BONE, WATER]
    dens_of = np.array([0.001205, 1.03, 0.95, 0.26, 1.92, 1.0], dtype=np.float32)
    mat = np.zeros(shape, dtype=np.uint8)
    xy = (x[None, :] / 160.0) ** 2 + (y[:, None] / 110.0) ** 2
    body = xy <= 1.0
    inner = (x[None, :] / 154.0) ** 2 + (y[:, None] / 104.0) ** 2 <= 1.0
    slice_mat = np.zeros(xy.shape, dtype=np.uint8)
    slice_mat[body] = 2
    slice_mat[inner] = 1
    mat[:] = slice_mat[None]
    # ribs: bone shell between two ellipses, upper half of the body length
    shell = ((x[None, :] / 140.0) ** 2 + (y[:, None] / 92.0) ** 2 <= 1.0) & ((x[None, :] / 132.0) ** 2 + (y[:, None] / 85.0) ** 2 > 1.0)
    for k in np.nonzero(z > -40)[0]:
        mat[k][shell] = 4
    for cx in (-65.0, 65.0):  # lungs
        _paint(mat, z, y, x, (cx, -5.0, 40.0), (50.0, 65.0, 110.0), 3)
    spine = ((x[None, :] / 18.0) ** 2 + ((y[:, None] - 70.0) / 18.0) ** 2 <= 1.0)
    mat[:, spine] = 4
    _paint(mat, z, y, x, (40.0, 20.0, -70.0), (45.0, 35.0, 40.0), 5)  # stomach (water)
    rng = np.random.default_rng(seed)
    dens = dens_of[mat]
    dens *= rng.uniform(0.98, 1.02, size=shape).astype(np.float32)
    spec, exposures, per = ct_spiral_spec(histories_total)
    return Workload("ct_spiral_120kVp_512x512x300", mat, dens.astype(np.float32), spacing, names, spec, per * exposures,
                    "CT spiral 120 kVp, 512x512x300 x 0.8 mm anthropomorphic ellipsoids, 3600 exposures",
                    {"exposures": exposures, "histories_per_exposure": per})


def whole_body(histories_total=10_000_000_000, seed=13, shape=(600, 1024, 1024), spacing=(0.7, 0.7, 0.7)):
    z, y, x = _grid(shape, spacing)
    names = [AIR, SOFT, ADIPOSE, LUNG, BONE, WATER, "Muscle, Skeletal (ICRP)", "Breast Tissue (ICRU)"]
    dens_of = np.array([0.001205, 1.03, 0.95, 0.26, 1.92, 1.0, 1.04, 1.02], dtype=np.float32)
    mat = np.zeros(shape, dtype=np.uint8)
    body = (x[None, :] / 170.0) ** 2 + (y[:, None] / 115.0) ** 2 <= 1.0
    inner = (x[None, :] / 163.0) ** 2 + (y[:, None] / 108.0) ** 2 <= 1.0
    muscle = (x[None, :] / 150.0) ** 2 + (y[:, None] / 96.0) ** 2 <= 1.0
    s = np.zeros(body.shape, dtype=np.uint8)
    s[body] = 2
    s[inner] = 6
    s[muscle] = 1
    mat[:] = s[None]
    rng = np.random.default_rng(seed)
    for cx in (-70.0, 70.0):
        _paint(mat, z, y, x, (cx, -5.0, 60.0), (55.0, 70.0, 120.0), 3)
    spine = ((x[None, :] / 20.0) ** 2 + ((y[:, None] - 75.0) / 20.0) ** 2 <= 1.0)
    mat[:, spine] = 4
    _paint(mat, z, y, x, (45.0, 25.0, -60.0), (45.0, 35.0, 40.0), 5)
    for cx in (-60.0, 60.0):
        _paint(mat, z, y, x, (cx, -95.0, 90.0), (40.0, 25.0, 40.0), 7)
    dens = dens_of[mat] * rng.uniform(0.98, 1.02, size=shape).astype(np.float32)
    per = int(round(histories_total / 360))
    ang, w = bowtie_profile()
    spec = source_spec(capi.SOURCE_CT_AXIAL, tube_voltage=120.0, anode_angle_deg=12.0, al_mm=6.0, sdd=1190.0, fov=500.0, collimation=38.4,
                       exposure_angle_step_deg=1.0, histories_per_exposure=per, bowtie_angles=ang, bowtie_weights=w, heel_effect=1)
    return Workload("whole_body_ct_axial_1024x1024x600", mat, dens.astype(np.float32), spacing, names, spec, per * 360,
                    "CT axial 120 kVp, 1024x1024x600 x 0.7 mm, 360 exposures")


def serving_requests(n=256, rate=20.0, seed=42, max_side=256, min_hist=1e6, max_hist=1e8):
    """Request stream of the serving configuration: arrival times (s) and specs."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    reqs = []
    for i in range(n):
        side = int(rng.integers(32, max_side + 1))
        sp = float(rng.uniform(1.0, 2.0))
        kvp = float(rng.uniform(80, 140))
        hist = int(round(10 ** rng.uniform(math.log10(min_hist), math.log10(max_hist))))
        dx = bool(rng.random() < 0.5)
        reqs.append({"arrival": float(t[i]), "side": side, "spacing": sp, "kvp": kvp, "histories": hist, "dx": dx, "seed": int(rng.integers(1, 2**31))})
    return reqs


def cube_phantom(side, spacing, seed):
    """Water cube with a seeded bone sphere (serving requests)."""
    rng = np.random.default_rng(seed)
    shape = (side, side, side)
    z, y, x = _grid(shape, (spacing,) * 3)
    mat = np.zeros(shape, dtype=np.uint8)
    half = side * spacing / 2
    r = rng.uniform(0.1, 0.3) * half
    c = rng.uniform(-half + r, half - r, 3)
    _paint(mat, z, y, x, c, (r, r, r), 1)
    dens = np.array([1.0, 1.92], dtype=np.float32)[mat]
    return mat, dens, [WATER, BONE]


def request_spec(req):
    """Source of a serving request: DX from 1 m or one CT axial rotation (1 deg steps)."""
    side_mm = req["side"] * req["spacing"]
    if req["dx"]:
        exposures = 100
        per = max(1, req["histories"] // exposures)
        return source_spec(capi.SOURCE_DX, tube_voltage=req["kvp"], al_mm=2.5, sdd=1000.0, field_size=(side_mm, side_mm),
                           histories_per_exposure=per, total_exposures=exposures, heel_effect=1)
    per = max(1, req["histories"] // 360)
    ang, w = bowtie_profile()
    return source_spec(capi.SOURCE_CT_AXIAL, tube_voltage=req["kvp"], al_mm=6.0, sdd=1190.0, fov=500.0, collimation=min(38.4, side_mm),
                       exposure_angle_step_deg=1.0, histories_per_exposure=per, bowtie_angles=ang, bowtie_weights=w, heel_effect=1)
