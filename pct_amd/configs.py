"""Scan and reconstruction presets.

Each preset gives the reconstruction parameters (the reference's
pCT_config.h / pCT_data.h set, include/pct.h pct_params) and the synthetic
scan that feeds them (include/pct_synth_api.h pct_synth_cfg).

  mini : 40x40x8 voxels of 5x5x2.5 mm, 20 000 histories over 90 angles
         (small enough for the pure-Python checks and committed fixtures)
  cfg1 : 100x100x16 voxels of 2x2x2.5 mm, 2 M histories, lambda 1e-4
         (the reference's CPU-runnable case)
  cfg2 : 200x200x32 voxels of 1x1x2.5 mm (the reference defaults), 20 M
         histories, lambda 1e-4 (the single-GPU benchmark workload)

Common synthetic geometry: tracker planes at u = -25, -15, +15, +25 cm, entry
hits uniform in |t| < 11 cm and |v| < 2 cm at u = -15 cm, cone beam from a
point source at SOURCE_RADIUS, 200 MeV protons (beta c p = 364.86 MeV),
phantom = 20 cm diameter water cylinder (RSP 1.0) with four 2 cm diameter
inserts (RSP 0.3, 1.1, 1.3, 1.6) centred 5 cm off axis at 45/135/225/315 deg,
1 % Gaussian WEPL straggling, 90 angles at 4 deg, seed 0x5EED2026 + config id.
"""
import math
from dataclasses import dataclass

from . import _lib

BETA_CP_200MEV = 364.86   # MeV: T = 200 MeV, m_p = 938.272 MeV -> p = 644.4 MeV/c, beta = 0.5662

INSERT_RSP = (0.3, 1.1, 1.3, 1.6)
INSERT_ANGLES_DEG = (45.0, 135.0, 225.0, 315.0)


@dataclass
class Preset:
    name: str
    columns: int
    rows: int
    slices: int
    slice_thickness: float
    n_histories: int
    lam: float = 1e-4
    block_size: int = 3200
    iterations: int = 12
    v_half: float = 2.0
    config_id: int = 0


PRESETS = {
    "mini": Preset("mini", 40, 40, 8, 0.25, 20_000, v_half=1.0, config_id=9),
    "cfg1": Preset("cfg1", 100, 100, 16, 0.25, 2_000_000, config_id=1),
    "cfg2": Preset("cfg2", 200, 200, 32, 0.25, 20_000_000, config_id=2),
}

SEED_BASE = 0x5EED2026


def params(preset: str, **overrides) -> _lib.PctParams:
    """pct_params for a preset (reference defaults for everything else)."""
    pr = PRESETS[preset]
    p = _lib.default_params()
    p.columns, p.rows, p.slices = pr.columns, pr.rows, pr.slices
    p.slice_thickness = pr.slice_thickness
    p.recon_cyl_height = pr.slices * pr.slice_thickness
    p.lambda_ = pr.lam
    p.drop_block_size = pr.block_size
    p.iterations = pr.iterations
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def synth_cfg(preset: str, n_histories: int = None, seed: int = None) -> _lib.PctSynthCfg:
    pr = PRESETS[preset]
    c = _lib.PctSynthCfg()
    c.seed = SEED_BASE + pr.config_id if seed is None else seed
    c.n_histories = pr.n_histories if n_histories is None else n_histories
    c.n_angles = 90
    c.angle_step_deg = 4.0
    c.u_in_1, c.u_in_2, c.u_out_1, c.u_out_2 = -25.0, -15.0, 15.0, 25.0
    c.t_half = 11.0
    c.v_half = pr.v_half
    c.source_radius = 260.7
    c.beta_cp = BETA_CP_200MEV
    c.x0 = 36.08
    c.straggle = 0.01
    cyl = [(0.0, 0.0, 10.0, 1.0)]
    for rsp, ang in zip(INSERT_RSP, INSERT_ANGLES_DEG):
        a = math.radians(ang)
        cyl.append((5.0 * math.cos(a), 5.0 * math.sin(a), 1.0, rsp))
    c.n_cyl = len(cyl)
    for k, (x, y, r, rsp) in enumerate(cyl):
        c.cyl_x[k], c.cyl_y[k], c.cyl_r[k], c.cyl_rsp[k] = x, y, r, rsp
    return c


def phantom_rsp(cfg: _lib.PctSynthCfg, p: _lib.PctParams):
    """Ground-truth RSP at voxel centres (numpy, [slices, rows, columns])."""
    import numpy as np
    C, R, S = p.columns, p.rows, p.slices
    w = 2 * p.recon_cyl_radius / C
    h = 2 * p.recon_cyl_radius / R
    xs = -p.recon_cyl_radius + (np.arange(C) + 0.5) * w
    ys = p.recon_cyl_radius - (np.arange(R) + 0.5) * h
    X, Y = np.meshgrid(xs, ys)
    img = np.zeros((R, C))
    for k in range(cfg.n_cyl):
        inside = (X - cfg.cyl_x[k]) ** 2 + (Y - cfg.cyl_y[k]) ** 2 < cfg.cyl_r[k] ** 2
        img[inside] = cfg.cyl_rsp[k]
    return np.broadcast_to(img, (S, R, C)).copy()


# ---------------------------------------------------------------------------
# Batched form: 12 independent scans of uneven size in one DROP iteration
# ---------------------------------------------------------------------------
BATCH_SCANS = 12
BATCH_GRIDS_XY = (64, 96, 128, 160, 192, 224, 256)
BATCH_SLICES = (8, 16, 24, 32, 40, 48)
BATCH_HISTORIES = (500_000, 30_000_000)      # log-uniform range
BATCH_BLOCK_SIZES = (800, 1600, 3200, 6400)  # the tuner's candidates
BATCH_RSP = (0.2, 1.8)                       # uniform range of every body's RSP


@dataclass
class BatchScan:
    name: str
    params: object        # pct_params
    synth: object         # pct_synth_cfg


def _random_ellipsoids(rng, height):
    """body + 3-8 non-overlapping inserts inside it (rejection sampling)"""
    body = dict(x=rng.uniform(-0.5, 0.5), y=rng.uniform(-0.5, 0.5), z=0.0, a=rng.uniform(7.0, 9.5),
                b=rng.uniform(6.0, 9.5), c=height / 2 + 2.0, phi=rng.uniform(0, math.pi), rsp=rng.uniform(*BATCH_RSP))
    out = [body]
    want = int(rng.integers(3, 9))
    tries = 0
    while len(out) < 1 + want and tries < 2000:
        tries += 1
        a, b, c = rng.uniform(0.5, 2.5), rng.uniform(0.5, 2.5), rng.uniform(0.5, max(0.6, height / 2))
        r = max(a, b)
        ang, rad = rng.uniform(0, 2 * math.pi), rng.uniform(0, min(body["a"], body["b"]) - r - 0.3)
        if rad < 0:
            continue
        e = dict(x=body["x"] + rad * math.cos(ang), y=body["y"] + rad * math.sin(ang), z=rng.uniform(-0.5, 0.5),
                 a=a, b=b, c=c, phi=rng.uniform(0, math.pi), rsp=rng.uniform(*BATCH_RSP))
        if all(math.hypot(e["x"] - o["x"], e["y"] - o["y"]) > r + max(o["a"], o["b"]) + 0.2 for o in out[1:]):
            out.append(e)
    return out


def batch_scans(seed: int = SEED_BASE + 4, scale: float = 1.0):
    """The batched configuration: 12 scans, grids from 64x64x8 to 256x256x48
    (the first and last scan take the two ends), histories log-uniform in
    [0.5 M, 30 M] (times `scale`), random ellipsoid phantoms with RSP in
    U[0.2, 1.8], 90 angles, max_intersections 1024 (paths up to ~450 voxels)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for k in range(BATCH_SCANS):
        if k == 0:
            cxy, s = 64, 8
        elif k == BATCH_SCANS - 1:
            cxy, s = 256, 48
        else:
            cxy, s = int(rng.choice(BATCH_GRIDS_XY)), int(rng.choice(BATCH_SLICES))
        lo, hi = math.log(BATCH_HISTORIES[0]), math.log(BATCH_HISTORIES[1])
        n = int(round(math.exp(rng.uniform(lo, hi)) * scale))
        p = _lib.default_params()
        p.columns = p.rows = cxy
        p.slices = s
        p.slice_thickness = 0.25
        p.recon_cyl_height = s * 0.25
        p.lambda_ = 1e-4
        p.drop_block_size = 3200
        p.iterations = 12
        p.max_intersections = 1024
        c = _lib.PctSynthCfg()
        c.seed = seed * 1000 + k
        c.n_histories = n
        c.n_angles = 90
        c.angle_step_deg = 4.0
        c.u_in_1, c.u_in_2, c.u_out_1, c.u_out_2 = -25.0, -15.0, 15.0, 25.0
        c.t_half = 11.0
        c.v_half = max(0.5, p.recon_cyl_height / 2 - 0.5)
        c.source_radius = 260.7
        c.beta_cp = BETA_CP_200MEV
        c.x0 = 36.08
        c.straggle = 0.01
        c.n_cyl = 0
        ells = _random_ellipsoids(rng, p.recon_cyl_height)
        c.n_ell = len(ells)
        for j, e in enumerate(ells):
            for f in ("x", "y", "z", "a", "b", "c", "phi", "rsp"):
                getattr(c, "ell_" + f)[j] = e[f]
        out.append(BatchScan(f"scan{k:02d}_{cxy}x{cxy}x{s}_{n}", p, c))
    return out


# ---------------------------------------------------------------------------
# Learned-path form: the large evaluation scan
# ---------------------------------------------------------------------------
LEARNED_TRAIN_PATHS = 10_000_000
LEARNED_TRAIN_SCANS = 8
LEARNED_TRAIN_HISTORIES_PER_SCAN = 1_600_000


def learned_large(n_histories: int = 50_000_000, seed: int = SEED_BASE + 6):
    """256x256x48 voxels (0.78 x 0.78 x 2.5 mm), n_histories over 90 angles, a
    random ellipsoid phantom (RSP in U[0.2, 1.8]) of its own seed,
    max_intersections 1024: the second evaluation case of the learned-path form"""
    import numpy as np
    rng = np.random.default_rng(seed)
    p = _lib.default_params()
    p.columns = p.rows = 256
    p.slices = 48
    p.slice_thickness = 0.25
    p.recon_cyl_height = 48 * 0.25
    p.lambda_ = 1e-4
    p.drop_block_size = 3200
    p.iterations = 12
    p.max_intersections = 1024
    c = _lib.PctSynthCfg()
    c.seed = seed
    c.n_histories = n_histories
    c.n_angles = 90
    c.angle_step_deg = 4.0
    c.u_in_1, c.u_in_2, c.u_out_1, c.u_out_2 = -25.0, -15.0, 15.0, 25.0
    c.t_half = 11.0
    c.v_half = p.recon_cyl_height / 2 - 0.5
    c.source_radius = 260.7
    c.beta_cp = BETA_CP_200MEV
    c.x0 = 36.08
    c.straggle = 0.01
    c.n_cyl = 0
    ells = _random_ellipsoids(rng, p.recon_cyl_height)
    c.n_ell = len(ells)
    for j, e in enumerate(ells):
        for f in ("x", "y", "z", "a", "b", "c", "phi", "rsp"):
            getattr(c, "ell_" + f)[j] = e[f]
    return BatchScan(f"large_256x256x48_{n_histories}", p, c)
