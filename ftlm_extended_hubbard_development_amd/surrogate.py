"""Learned surrogate of the FTLM thermodynamics, with FTLM as judge and fallback.

The reference computes n(mu) and the compressibility for one parameter set
per run of ftlm_n_vs_mu (src/main.cpp:1459-1839); a scan over twists or
model parameters is one full sweep per point. This module adds:

* ``Surrogate``: the MLP of libftlm_hip.so (``ftlm_mlp_*`` in include/ftlm.h)
  mapping (U, V, t', phix, phiy, beta, mu) to (n, compressibility), trained
  and evaluated on the GPU by the library's bf16-MFMA kernels (fp32
  accumulation);
* ``generate``: training data from the engine's own FTLM sweeps (the drop-in
  ``ftlm_n_vs_mu`` executable on the GPU) at seeded parameter draws;
* ``scan``: a dense twist scan filled by the surrogate, where FTLM runs on a
  coarse sub-grid of judge twists and replaces the surrogate on every twist
  whose neighbouring judges show an error above the tolerance
  (|dn| > 1e-3 by default), so FTLM stays the judge and the fallback.

Command line::

    python -m ftlm_extended_hubbard_development_amd.surrogate generate --lattice 3x3 --draws 48 --out d33.npz
    python -m ftlm_extended_hubbard_development_amd.surrogate train d42.npz d33.npz --out model.npz
    python -m ftlm_extended_hubbard_development_amd.surrogate scan model.npz --lattice 3x3 --grid 16 ...
    python -m ftlm_extended_hubbard_development_amd.surrogate gain pretrained:4x2 --lattice 4x2 --u 8 --exact

A model argument ``pretrained:LATTICE`` names the parameters shipped in
``data/`` (``PRETRAINED``).
"""
import argparse
import ctypes as C
import glob
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

from . import FTLM_EXE, FtlmError, FtlmMlpConfig, _fp, _dp, _p, load_library

# parameter ranges of the training draws
RANGES = {"u": (4.0, 10.0), "v": (0.0, 2.0), "tp": (-0.3, 0.3), "phix": (0.0, 1.0), "phiy": (0.0, 1.0),
          "beta": (2.0, 20.0), "mu": (-5.0, 15.0)}
INPUTS = ("u", "v", "tp", "phix", "phiy", "beta", "mu")
TOL_N = 1e-3  # |dn| above which FTLM replaces the surrogate

# trained parameters shipped with the package (the command that made each
# is in its "meta" entry)
PRETRAINED = {"4x2": os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "surrogate_4x2.npz")}


def pretrained(lattice, device=0, batch=4096):
    """The shipped surrogate of a lattice ("4x2"), loaded on the GPU."""
    if lattice not in PRETRAINED:
        raise ValueError(f"no pretrained surrogate for {lattice!r}; have {sorted(PRETRAINED)}")
    return Surrogate.load(PRETRAINED[lattice], device=device, batch=batch)


def layout():
    """(weight offsets, bias offsets, out widths, in widths) of the flat
    parameter vector."""
    lib = load_library()
    n = 5
    wo = np.zeros(n, np.int64)
    bo = np.zeros(n, np.int64)
    no = np.zeros(n, np.int32)
    ni = np.zeros(n, np.int32)
    lib.ftlm_mlp_layout(_p(wo, C.POINTER(C.c_int64)), _p(bo, C.POINTER(C.c_int64)), _p(no, C.POINTER(C.c_int32)),
                        _p(ni, C.POINTER(C.c_int32)))
    return wo, bo, no, ni


def param_count():
    return int(load_library().ftlm_mlp_param_count())


# inputs actually used by the first layer and outputs of the last one; the
# rest of the padded widths stays zero
USED_FEATURES = 16
USED_OUTPUTS = 2


def init_params(seed=0):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases (fan_in =
    the used inputs), padding rows/columns zero."""
    rng = np.random.default_rng(seed)
    wo, bo, no, ni = layout()
    p = np.zeros(param_count(), np.float32)
    for l in range(5):
        fan_in = USED_FEATURES if l == 0 else int(ni[l])
        a = 1.0 / np.sqrt(fan_in)
        W = rng.uniform(-a, a, size=(int(no[l]), int(ni[l])))
        b = rng.uniform(-a, a, size=int(no[l]))
        if l == 0:
            W[:, USED_FEATURES:] = 0.0
        if l == 4:
            W[USED_OUTPUTS:, :] = 0.0
            b[USED_OUTPUTS:] = 0.0
        p[wo[l]:wo[l] + W.size] = W.ravel()
        p[bo[l]:bo[l] + b.size] = b
    return p


def encode(x):
    """Feature rows (n, 32) float32 of input rows (n, 7) (host code of the
    library)."""
    x = np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1, 7))
    f = np.zeros((x.shape[0], 32), np.float32)
    rc = load_library().ftlm_mlp_encode(x.shape[0], _p(x, _dp), _p(f, _fp))
    if rc != 0:
        raise FtlmError(load_library().ftlm_last_error().decode())
    return f


class Surrogate:
    """The MLP on one GPU."""

    def __init__(self, y_mean=(1.0, 0.0), y_scale=(1.0, 1.0), batch=4096, lr=1e-3, beta1=0.9, beta2=0.999,
                 eps=1e-8, params=None, seed=0, device=0):
        self.lib = load_library()
        self.cfg = FtlmMlpConfig(batch, lr, beta1, beta2, eps, (C.c_double * 2)(*y_mean),
                                 (C.c_double * 2)(*y_scale))
        h = C.c_void_p()
        self._check(self.lib.ftlm_mlp_create(C.byref(self.cfg), device, C.byref(h)))
        self.h = h
        self.P = param_count()
        self.set_params(init_params(seed) if params is None else params)

    def _check(self, rc):
        if rc != 0:
            raise FtlmError(f"ftlm error {rc}: {self.lib.ftlm_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.ftlm_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def y_mean(self):
        return tuple(self.cfg.y_mean)

    @property
    def y_scale(self):
        return tuple(self.cfg.y_scale)

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.P
        self._check(self.lib.ftlm_mlp_set_params(self.h, _p(p, _fp)))

    def params(self):
        p = np.zeros(self.P, np.float32)
        self._check(self.lib.ftlm_mlp_get_params(self.h, _p(p, _fp)))
        return p

    def set_lr(self, lr):
        self._check(self.lib.ftlm_mlp_set_lr(self.h, lr))

    def set_data(self, x, y):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 7)
        y = np.ascontiguousarray(y, np.float64).reshape(-1, 2)
        assert x.shape[0] == y.shape[0]
        self._check(self.lib.ftlm_mlp_set_data(self.h, x.shape[0], _p(x, _dp), _p(y, _dp)))

    def train(self, steps, seed=0):
        loss = np.zeros(steps, np.float32)
        self._check(self.lib.ftlm_mlp_train(self.h, steps, seed, _p(loss, _fp)))
        return loss

    def gradients(self, x, y):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 7)
        y = np.ascontiguousarray(y, np.float64).reshape(-1, 2)
        g = np.zeros(self.P, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self.lib.ftlm_mlp_gradients(self.h, x.shape[0], _p(x, _dp), _p(y, _dp), _p(g, _fp),
                                                _p(loss, _fp)))
        return g, float(loss[0])

    def step(self, x, y):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 7)
        y = np.ascontiguousarray(y, np.float64).reshape(-1, 2)
        loss = np.zeros(1, np.float32)
        self._check(self.lib.ftlm_mlp_step(self.h, x.shape[0], _p(x, _dp), _p(y, _dp), _p(loss, _fp)))
        return float(loss[0])

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 7)
        y = np.zeros((x.shape[0], 2))
        if x.shape[0]:
            self._check(self.lib.ftlm_mlp_predict(self.h, x.shape[0], _p(x, _dp), _p(y, _dp)))
        return y

    def gemm(self, a, b):
        """C = bf16(a) bf16(b)^T through the network's MFMA kernel (test hook)."""
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        c = np.zeros((a.shape[0], b.shape[0]), np.float32)
        self._check(self.lib.ftlm_mlp_gemm(self.h, a.shape[0], b.shape[0], a.shape[1], _p(a, _fp), _p(b, _fp),
                                           _p(c, _fp)))
        return c

    def save(self, path, meta=None):
        np.savez(path, params=self.params(), y_mean=np.array(self.y_mean), y_scale=np.array(self.y_scale),
                 meta=np.array(json.dumps(meta or {})))

    @classmethod
    def load(cls, path, device=0, batch=4096):
        d = np.load(path, allow_pickle=False)
        return cls(y_mean=tuple(d["y_mean"]), y_scale=tuple(d["y_scale"]), params=d["params"], batch=batch,
                   device=device)


# the compressibility is scaled by this many standard deviations more than n
# in the loss: its peaks (kappa = beta * charge fluctuation) would otherwise
# dominate, while the FTLM gate is on n
KAPPA_DOWNWEIGHT = 4.0


def train_schedule(s, steps, lr, lr_final, chunks, seed, log=None):
    """Adam with a geometric learning-rate decay from lr to lr_final over
    `chunks` equal parts; returns the loss history."""
    hist = []
    per = max(1, steps // chunks)
    done = 0
    for c in range(chunks):
        rate = lr * (lr_final / lr) ** (c / max(1, chunks - 1))
        s.set_lr(rate)
        n = per if c < chunks - 1 else steps - done
        if n <= 0:
            break
        hist.append(s.train(n, seed=seed * 1000003 + c))
        done += n
        if log:
            log(f"chunk {c}: lr {rate:.2e} loss {float(hist[-1][-min(50, n):].mean()):.3e}")
    return np.concatenate(hist)


def fit(x, y, steps=6000, batch=4096, lr=2e-3, lr_final=2e-5, chunks=12, seed=0, device=0, log=None):
    """Train a new surrogate on (x, y): targets normalised by their mean and
    standard deviation (the compressibility's scaled KAPPA_DOWNWEIGHT times
    more), Adam with a geometric learning-rate decay. Returns (surrogate,
    loss history)."""
    x = np.asarray(x, np.float64).reshape(-1, 7)
    y = np.asarray(y, np.float64).reshape(-1, 2)
    mean = y.mean(axis=0)
    scale = np.maximum(y.std(axis=0), 1e-6) * np.array([1.0, KAPPA_DOWNWEIGHT])
    s = Surrogate(y_mean=tuple(mean), y_scale=tuple(scale), batch=batch, lr=lr, seed=seed, device=device)
    s.set_data(x, y)
    return s, train_schedule(s, steps, lr, lr_final, chunks, seed, log)


def finetune(base, x, y, steps=2000, batch=1024, lr=5e-4, lr_final=1e-5, chunks=4, seed=0, device=0):
    """A copy of `base` (same normalisation) trained further on (x, y)."""
    s = Surrogate(y_mean=base.y_mean, y_scale=base.y_scale, batch=batch, lr=lr, params=base.params(),
                  device=device)
    s.set_data(x, y)
    return s, train_schedule(s, steps, lr, lr_final, chunks, seed)


# ---------------------------------------------------------------- FTLM data
def parse_lattice(s):
    lx, ly = (int(t) for t in s.lower().split("x"))
    return lx, ly


def _read_multi_beta_csv(path):
    """beta, mu, n, compressibility of a multi-beta CSV (main.cpp:1427-1455)."""
    with open(path) as f:
        head = f.readline().strip().split(",")
    data = np.loadtxt(path, delimiter=",", skiprows=1, ndmin=2)
    col = {k: i for i, k in enumerate(head)}
    return data[:, col["beta"]], data[:, col["mu"]], data[:, col["n"]], data[:, col["compressibility"]]


def ftlm_twists(lattice, u, v, tp, twists, betas, mu_min=-5.0, mu_max=15.0, mu_count=81, samples=32, steps=80,
                seed=12345, exe=FTLM_EXE, chunk=48, log=None, extra=()):
    """Run the drop-in FTLM executable (GPU) for one model and a list of
    twists, `chunk` twists per process. Returns (rows (n, 7) inputs, (n, 2)
    outputs, wall seconds)."""
    xs, ys, wall = [], [], 0.0
    for c0 in range(0, len(twists), chunk):
        x, y, w = _ftlm_twists_once(lattice, u, v, tp, twists[c0:c0 + chunk], betas, mu_min, mu_max, mu_count,
                                    samples, steps, seed, exe, extra)
        xs.append(x)
        ys.append(y)
        wall += w
        if log:
            log(f"  FTLM {lattice} twists {c0}..{min(len(twists), c0 + chunk) - 1} of {len(twists)}: {w:.2f}s")
    return np.concatenate(xs), np.concatenate(ys), wall


def _ftlm_twists_once(lattice, u, v, tp, twists, betas, mu_min, mu_max, mu_count, samples, steps, seed, exe,
                      extra, workdir=None):
    lx, ly = parse_lattice(lattice)
    betas = [float(b) for b in betas]
    nb = len(betas)
    # the multi-beta schema needs two betas; a repeated beta repeats its rows
    cli_betas = betas * 2 if nb < 2 else betas
    tmp = workdir or tempfile.mkdtemp(prefix="ftlm_sur_")
    out = os.path.join(tmp, "scan.csv")
    tw = ",".join(f"{float(px)!r}:{float(py)!r}" for px, py in twists)
    cmd = [exe, "--lx", str(lx), "--ly", str(ly), "--u", repr(float(u)), "--v", repr(float(v)), "--tp",
           repr(float(tp)), "--beta-list", ",".join(repr(b) for b in cli_betas), "--mu-min", repr(float(mu_min)),
           "--mu-max", repr(float(mu_max)), "--mu-count", str(mu_count), "--samples", str(samples),
           "--lanczos-steps", str(steps), "--seed", str(seed), "--twist-list", tw, "--output", out, *extra]
    t0 = time.perf_counter()
    r = subprocess.run(cmd, capture_output=True, text=True)
    wall = time.perf_counter() - t0
    if r.returncode != 0:
        raise FtlmError(f"ftlm_n_vs_mu failed ({r.returncode}): {r.stderr.strip()[-500:]}")
    xs, ys = [], []
    bcol = np.repeat(np.array(betas), mu_count)  # rows run over betas (list order), then mu
    for i, (px, py) in enumerate(twists):
        path = os.path.join(tmp, f"scan.twist{i}.csv")
        _, m, n, k = _read_multi_beta_csv(path)
        if m.size != len(cli_betas) * mu_count:
            raise FtlmError(f"unexpected row count {m.size} in {path}")
        keep = slice(0, nb * mu_count)
        m, n, k = m[keep], n[keep], k[keep]
        rows = m.size
        xs.append(np.stack([np.full(rows, u), np.full(rows, v), np.full(rows, tp), np.full(rows, float(px)),
                            np.full(rows, float(py)), bcol, m], axis=1))
        ys.append(np.stack([n, k], axis=1))
        os.remove(path)
    if workdir is None:
        for f in glob.glob(os.path.join(tmp, "*")):
            os.remove(f)
        os.rmdir(tmp)
    return np.concatenate(xs), np.concatenate(ys), wall


def draw_models(n, seed):
    """Seeded draws of (U, V, t') in their ranges."""
    rng = np.random.default_rng(seed)
    return [(rng.uniform(*RANGES["u"]), rng.uniform(*RANGES["v"]), rng.uniform(*RANGES["tp"])) for _ in range(n)]


def generate(lattice, draws, twists_per_draw=8, betas=None, mu_count=81, samples=32, steps=80, seed=1, log=None):
    """Training data from the engine's own FTLM sweeps: `draws` seeded (U, V,
    t') draws, each at `twists_per_draw` random twists, all betas and mu of one
    multi-beta sweep. Returns (x (n, 7), y (n, 2), seconds of FTLM)."""
    if betas is None:
        betas = list(np.geomspace(RANGES["beta"][0], RANGES["beta"][1], 8))
    rng = np.random.default_rng(seed + 7919)
    xs, ys, secs = [], [], 0.0
    for d, (u, v, tp) in enumerate(draw_models(draws, seed)):
        twists = [(float(rng.uniform(0, 1)), float(rng.uniform(0, 1))) for _ in range(twists_per_draw)]
        x, y, wall = ftlm_twists(lattice, u, v, tp, twists, betas, *RANGES["mu"], mu_count=mu_count,
                                 samples=samples, steps=steps)
        xs.append(x)
        ys.append(y)
        secs += wall
        if log:
            log(f"{lattice} draw {d}: U={u:.3f} V={v:.3f} tp={tp:.3f} {len(twists)} twists {wall:.2f}s")
    return np.concatenate(xs), np.concatenate(ys), secs


# -------------------------------------------------------------------- scans
def twist_grid(n):
    return [(ix / n, iy / n) for iy in range(n) for ix in range(n)]


def scan(sur, lattice, u, v, tp, grid, betas, mu_min=-5.0, mu_max=15.0, mu_count=81, judge_stride=4, tol=TOL_N,
         samples=32, steps=80, seed=12345, finetune_steps=2000, log=None):
    """Dense twist scan (grid x grid twists, all betas and mu) filled by the
    surrogate, with FTLM as the judge and the fallback.

    The twist grid is cut into cells of judge_stride x judge_stride twists.
    FTLM runs on two coarse sub-grids: the cell corners ("fit" twists) and the
    cell centres ("check" twists). With finetune_steps > 0 a copy of the
    surrogate is first trained further on the fit twists' FTLM rows (the
    scan's own model parameters); the check twists are never trained on. A
    cell keeps the surrogate's values only when its check twist agrees with
    FTLM within `tol` in n at every (beta, mu); every other twist of a failing
    cell is computed by FTLM. Fit and check twists always take FTLM's values.

    Returns dict: x (rows of inputs), y (served values), source (0 surrogate,
    1 FTLM) per row and per twist, check errors, timings."""
    s = judge_stride
    if grid % s or s < 2:
        raise ValueError("grid must be a multiple of judge_stride >= 2")
    tw = twist_grid(grid)
    mu = np.linspace(mu_min, mu_max, mu_count)
    betas = [float(b) for b in betas]
    nb, nm = len(betas), mu_count
    ftlm = dict(mu_min=mu_min, mu_max=mu_max, mu_count=mu_count, samples=samples, steps=steps, seed=seed, log=log)
    fit_idx = [iy * grid + ix for iy in range(0, grid, s) for ix in range(0, grid, s)]
    chk_idx = [iy * grid + ix for iy in range(s // 2, grid, s) for ix in range(s // 2, grid, s)]
    judges = fit_idx + chk_idx
    xj, yj, t_judge = ftlm_twists(lattice, u, v, tp, [tw[i] for i in judges], betas, **ftlm)
    yj = yj.reshape(len(judges), nb * nm, 2)
    t0 = time.perf_counter()
    model = sur
    if finetune_steps > 0:
        rows = len(fit_idx) * nb * nm
        model, _ = finetune(sur, xj[:rows], yj[:len(fit_idx)].reshape(-1, 2), steps=finetune_steps,
                            batch=min(1024, (rows // 64) * 64 or 64))
    xs = np.array([[u, v, tp, px, py, b, m] for (px, py) in tw for b in betas for m in mu])
    ys = model.predict(xs).reshape(len(tw), nb * nm, 2)
    if model is not sur:
        model.close()
    t_sur = time.perf_counter() - t0
    chk_err = np.abs(ys[chk_idx, :, 0] - yj[len(fit_idx):, :, 0]).max(axis=1)
    cell_ok = {}
    for i, e in zip(chk_idx, chk_err):
        cell_ok[((i % grid) // s, (i // grid) // s)] = e <= tol
    src = np.zeros(len(tw), np.int8)
    out = ys.copy()
    for k, i in enumerate(judges):
        out[i] = yj[k]
        src[i] = 1
    fall = [i for i in range(len(tw)) if not src[i] and not cell_ok[((i % grid) // s, (i // grid) // s)]]
    t_fall = 0.0
    if fall:
        _, yf, t_fall = ftlm_twists(lattice, u, v, tp, [tw[i] for i in fall], betas, **ftlm)
        yf = yf.reshape(len(fall), nb * nm, 2)
        for k, i in enumerate(fall):
            out[i] = yf[k]
            src[i] = 1
    return {"x": xs, "y": out.reshape(-1, 2), "y_surrogate": ys.reshape(-1, 2), "source": np.repeat(src, nb * nm),
            "twist_source": src, "fit": fit_idx, "check": chk_idx, "check_err": chk_err, "fallback": fall,
            "t_surrogate": t_sur, "t_judge": t_judge, "t_fallback": t_fall}


def evaluate(sur, x, y, tol=TOL_N):
    """Error statistics of the surrogate on held-out rows: |dn| quantiles,
    the fraction within `tol`, the maximum per beta, and |dkappa|."""
    pred = sur.predict(x)
    dn = np.abs(pred[:, 0] - y[:, 0])
    dk = np.abs(pred[:, 1] - y[:, 1])
    per_beta = {f"{b:.4g}": float(dn[x[:, 5] == b].max()) for b in np.unique(x[:, 5])}
    return {"rows": int(x.shape[0]), "dn_max": float(dn.max()), "dn_p50": float(np.quantile(dn, 0.5)),
            "dn_p90": float(np.quantile(dn, 0.9)), "dn_p99": float(np.quantile(dn, 0.99)),
            "frac_within_tol": float((dn <= tol).mean()), "dkappa_max": float(dk.max()),
            "dkappa_p50": float(np.quantile(dk, 0.5)), "dn_max_per_beta": per_beta}


def gain(sur, lattice, u, v, tp, grid, betas, judge_stride=4, tol=TOL_N, samples=32, steps=80, finetune_steps=2000,
         mu_min=-5.0, mu_max=15.0, mu_count=81, save=None, exact=False, low_samples=(), log=None):
    """What the surrogate gains on a dense twist scan: the judged scan
    (surrogate + FTLM judges + FTLM fallback) against FTLM on every twist
    (the method without the surrogate). Returns times, the number of twists
    the surrogate served and the error of the served values against the
    all-FTLM scan (judge values are FTLM's own, so only surrogate-served rows
    can differ). interp_check_err is the error of the plain mean of FTLM at
    each check twist's four cell corners, the part of the check error that
    comes from the cell size rather than from the network. With `save`, the
    scan's rows (all-FTLM, surrogate and served values) go to that .npz.
    With `exact`, every twist is also computed with every block diagonalised
    exactly, and both scans are measured against that: FTLM's own sampling
    error against the error of the judged scan at each tolerance (the gain
    at equal error), and, for each R in `low_samples`, FTLM on every twist
    with R random vectors (the cheaper scan without the network)."""
    if low_samples and not exact:
        raise ValueError("low_samples scans are measured against the exact answer: pass exact=True")
    mus = dict(mu_min=mu_min, mu_max=mu_max, mu_count=mu_count)
    r = scan(sur, lattice, u, v, tp, grid, betas, judge_stride=judge_stride, tol=tol, samples=samples, steps=steps,
             finetune_steps=finetune_steps, log=log, **mus)
    tw = twist_grid(grid)
    _, yf, t_full = ftlm_twists(lattice, u, v, tp, tw, betas, samples=samples, steps=steps, log=log, **mus)
    dn = np.abs(r["y"][:, 0] - yf[:, 0])
    sur_rows = r["source"] == 0
    t_scan = r["t_surrogate"] + r["t_judge"] + r["t_fallback"]
    # the same scan at other tolerances, from the measured pieces: cells
    # accepted at `t`, fallback time at the measured FTLM time per twist, the
    # error of what the surrogate then serves against FTLM (the all-FTLM run)
    nb_nm = yf.shape[0] // len(tw)
    t_twist = t_full / len(tw)
    ypred = r["y_surrogate"].reshape(len(tw), nb_nm, 2)
    dn_sur = np.abs(ypred[:, :, 0] - yf.reshape(len(tw), nb_nm, 2)[:, :, 0]).max(axis=1)
    judged = set(r["fit"]) | set(r["check"])
    s = judge_stride
    yex, t_exact = None, 0.0
    if exact:
        # every block diagonalised exactly: the scan's exact answer
        lx, ly = parse_lattice(lattice)
        _, yex, t_exact = ftlm_twists(lattice, u, v, tp, tw, betas, samples=samples, steps=steps, log=log,
                                      extra=("--exact-block-threshold", str(4 ** (lx * ly))), **mus)
        yex = yex.reshape(len(tw), nb_nm, 2)[:, :, 0]
    yfn = yf.reshape(len(tw), nb_nm, 2)[:, :, 0]
    curve = []
    for t in sorted({tol, 2e-3, 5e-3, 1e-2, 2e-2, 5e-2}):
        ok = {((i % grid) // s, (i // grid) // s) for i, e in zip(r["check"], r["check_err"]) if e <= t}
        served = [i for i in range(len(tw)) if i not in judged and ((i % grid) // s, (i // grid) // s) in ok]
        n_fall = len(tw) - len(judged) - len(served)
        t_est = r["t_surrogate"] + r["t_judge"] + n_fall * t_twist
        point = {"tol": t, "surrogate_twists": len(served), "speedup": t_full / t_est,
                 "dn_max_served": float(max((dn_sur[i] for i in served), default=0.0))}
        if yex is not None:
            yt = yfn.copy()
            yt[served] = ypred[served, :, 0]
            d = np.abs(yt - yex)
            point.update(dn_exact_max=float(d.max()), dn_exact_p50=float(np.median(d)))
        curve.append(point)
    # what the cell size alone allows: FTLM at the check twist against the
    # mean of FTLM at its four cell corners (the twist is periodic in 1)
    yg = yf.reshape(len(tw), nb_nm, 2)
    interp_err = []
    for i in r["check"]:
        x0, y0 = (i % grid) - s // 2, (i // grid) - s // 2
        corners = [((y0 + dy) % grid) * grid + (x0 + dx) % grid for dy in (0, s) for dx in (0, s)]
        interp_err.append(float(np.abs(yg[corners, :, 0].mean(axis=0) - yg[i, :, 0]).max()))
    if save:
        extra = {} if yex is None else {"n_exact": yex.ravel()}
        np.savez_compressed(save, x=r["x"], y_ftlm=yf, y_surrogate=r["y_surrogate"], y_served=r["y"],
                            twist_source=r["twist_source"], check=np.asarray(r["check"]),
                            check_err=r["check_err"], grid=grid, judge_stride=s, **extra)
    exact_out = {}
    if yex is not None:
        d = np.abs(yfn - yex)
        exact_out = {"t_exact": t_exact, "ftlm_dn_exact_max": float(d.max()),
                     "ftlm_dn_exact_p50": float(np.median(d))}
        # the method without the network made cheaper the usual way: fewer
        # random vectors on every twist
        for rl in low_samples:
            _, yl, tl = ftlm_twists(lattice, u, v, tp, tw, betas, samples=rl, steps=steps, log=log, **mus)
            d = np.abs(yl.reshape(len(tw), nb_nm, 2)[:, :, 0] - yex)
            exact_out.setdefault("ftlm_fewer_samples", []).append(
                {"samples": rl, "t": tl, "speedup": t_full / tl, "dn_exact_max": float(d.max()),
                 "dn_exact_p50": float(np.median(d))})
    return {"twists": len(tw), "judges": len(judged), "fallback": len(r["fallback"]),
            "surrogate_twists": int((r["twist_source"] == 0).sum()),
            "t_ftlm_all": t_full, "t_scan": t_scan, "t_surrogate": r["t_surrogate"], "t_judge": r["t_judge"],
            "t_fallback": r["t_fallback"], "speedup": t_full / t_scan,
            "dn_max_served": float(dn.max()),
            "dn_max_surrogate_rows": float(dn[sur_rows].max()) if sur_rows.any() else 0.0,
            "dn_p50_surrogate_rows": float(np.quantile(dn[sur_rows], 0.5)) if sur_rows.any() else 0.0,
            "check_err": [float(e) for e in r["check_err"]], "interp_check_err": interp_err,
            "tolerance_curve": curve, **exact_out}


# ---------------------------------------------------------------------- CLI
def _load_model(name):
    if name.startswith("pretrained:"):
        return pretrained(name.split(":", 1)[1])
    return Surrogate.load(name)


def _main(argv=None):
    ap = argparse.ArgumentParser(prog="surrogate", description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate")
    g.add_argument("--lattice", required=True)
    g.add_argument("--draws", type=int, default=32)
    g.add_argument("--twists", type=int, default=8)
    g.add_argument("--samples", type=int, default=32)
    g.add_argument("--steps", type=int, default=80)
    g.add_argument("--mu-count", type=int, default=81)
    g.add_argument("--seed", type=int, default=1)
    g.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("data", nargs="+")
    t.add_argument("--steps", type=int, default=6000)
    t.add_argument("--batch", type=int, default=4096)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--out", required=True)
    e = sub.add_parser("evaluate")
    e.add_argument("model")
    e.add_argument("data", nargs="+")
    for name in ("scan", "gain"):
        s = sub.add_parser(name)
        s.add_argument("model")
        s.add_argument("--lattice", required=True)
        s.add_argument("--u", type=float, required=True)
        s.add_argument("--v", type=float, default=0.0)
        s.add_argument("--tp", type=float, default=0.0)
        s.add_argument("--grid", type=int, default=16)
        s.add_argument("--judge-stride", type=int, default=4)
        s.add_argument("--betas", default="2,5,10")
        s.add_argument("--tol", type=float, default=TOL_N)
        s.add_argument("--samples", type=int, default=32)
        s.add_argument("--steps", type=int, default=80)
        s.add_argument("--finetune-steps", type=int, default=2000)
        if name == "scan":
            s.add_argument("--out", required=True)
        else:
            s.add_argument("--save", default=None)
            s.add_argument("--exact", action="store_true")
            s.add_argument("--low-samples", default="", help="with --exact: FTLM scans at these R, e.g. 4,8")
    a = ap.parse_args(argv)

    def log(m):
        print(m, flush=True)

    if a.cmd == "generate":
        x, y, secs = generate(a.lattice, a.draws, a.twists, mu_count=a.mu_count, samples=a.samples, steps=a.steps,
                              seed=a.seed, log=log)
        np.savez_compressed(a.out, x=x, y=y, ftlm_seconds=secs, lattice=a.lattice)
        log(f"wrote {a.out}: {x.shape[0]} rows, {secs:.1f}s of FTLM")
    elif a.cmd == "train":
        xs, ys = [], []
        for p in a.data:
            d = np.load(p, allow_pickle=False)
            xs.append(d["x"])
            ys.append(d["y"])
        x, y = np.concatenate(xs), np.concatenate(ys)
        t0 = time.perf_counter()
        sur, hist = fit(x, y, steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed, log=log)
        sur.save(a.out, meta={"data": a.data, "steps": a.steps, "train_s": time.perf_counter() - t0})
        log(f"wrote {a.out}: final loss {float(hist[-50:].mean()):.3e}")
    elif a.cmd == "evaluate":
        sur = _load_model(a.model)
        for p in a.data:
            d = np.load(p, allow_pickle=False)
            log(json.dumps({"data": p, **evaluate(sur, d["x"], d["y"])}))
    else:
        sur = _load_model(a.model)
        kw = dict(judge_stride=a.judge_stride, tol=a.tol, samples=a.samples, steps=a.steps,
                  finetune_steps=a.finetune_steps, log=log)
        betas = [float(b) for b in a.betas.split(",")]
        if a.cmd == "gain":
            log(json.dumps(gain(sur, a.lattice, a.u, a.v, a.tp, a.grid, betas, save=a.save, exact=a.exact,
                                  low_samples=[int(r) for r in a.low_samples.split(",") if r], **kw)))
        else:
            r = scan(sur, a.lattice, a.u, a.v, a.tp, a.grid, betas, **kw)
            np.savez(a.out, **{k: np.asarray(v) for k, v in r.items()})
            log(json.dumps({"twists": a.grid ** 2, "judges": len(r["fit"]) + len(r["check"]),
                            "fallback": len(r["fallback"]), "surrogate_twists": int((r["twist_source"] == 0).sum()),
                            "t_surrogate": r["t_surrogate"], "t_judge": r["t_judge"], "t_fallback": r["t_fallback"]}))
    return 0


if __name__ == "__main__":
    sys.exit(_main())
