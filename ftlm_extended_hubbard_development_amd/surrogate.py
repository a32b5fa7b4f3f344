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
USED_FEATURES = 12
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


def fit(x, y, steps=6000, batch=4096, lr=2e-3, lr_final=2e-5, chunks=12, seed=0, device=0, log=None):
    """Train a new surrogate on (x, y): targets normalised by their mean and
    standard deviation, Adam with a geometric learning-rate decay from lr to
    lr_final over `chunks` equal parts. Returns (surrogate, loss history)."""
    x = np.asarray(x, np.float64).reshape(-1, 7)
    y = np.asarray(y, np.float64).reshape(-1, 2)
    mean = y.mean(axis=0)
    scale = np.maximum(y.std(axis=0), 1e-6)
    s = Surrogate(y_mean=tuple(mean), y_scale=tuple(scale), batch=batch, lr=lr, seed=seed, device=device)
    s.set_data(x, y)
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
    return s, np.concatenate(hist)


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
                seed=12345, exe=FTLM_EXE, workdir=None, extra=()):
    """Run the drop-in FTLM executable (GPU) for one model and a list of
    twists in one process. Returns (rows (n, 7) inputs, (n, 2) outputs, wall s)."""
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
         samples=32, steps=80, seed=12345):
    """Dense twist scan (grid x grid twists, all betas and mu) filled by the
    surrogate. FTLM runs on the judge twists (every `judge_stride`-th twist in
    each direction); a twist keeps the surrogate's values only when every
    judge of its cell (the corners of the judge sub-grid cell holding it,
    periodic) agrees with FTLM within `tol` in n at every (beta, mu);
    otherwise FTLM is run for it (fallback). Judge twists take FTLM's values.

    Returns dict: x (rows of inputs), y (served values), source (0 surrogate,
    1 FTLM), judge_err (per judge twist), timings."""
    if grid % judge_stride:
        raise ValueError("grid must be a multiple of judge_stride")
    tw = twist_grid(grid)
    mu = np.linspace(mu_min, mu_max, mu_count)
    betas = [float(b) for b in betas]
    nb, nm = len(betas), mu_count
    # surrogate everywhere
    t0 = time.perf_counter()
    xs = np.array([[u, v, tp, px, py, b, m] for (px, py) in tw for b in betas for m in mu])
    ys = sur.predict(xs).reshape(len(tw), nb * nm, 2)
    t_sur = time.perf_counter() - t0
    # judges
    jidx = [iy * grid + ix for iy in range(0, grid, judge_stride) for ix in range(0, grid, judge_stride)]
    xj, yj, t_judge = ftlm_twists(lattice, u, v, tp, [tw[i] for i in jidx], betas, mu_min, mu_max, mu_count,
                                  samples, steps, seed)
    yj = yj.reshape(len(jidx), nb * nm, 2)
    jerr = np.abs(ys[jidx, :, 0] - yj[:, :, 0]).max(axis=1)
    jbad = {i: e > tol for i, e in zip(jidx, jerr)}
    src = np.zeros(len(tw), np.int8)
    out = ys.copy()
    for k, i in enumerate(jidx):
        out[i] = yj[k]
        src[i] = 1
    fall = []
    for i, (px, py) in enumerate(tw):
        if src[i]:
            continue
        ix, iy = i % grid, i // grid
        x0, y0 = ix - ix % judge_stride, iy - iy % judge_stride
        corners = [((y0 + dy) % grid) * grid + (x0 + dx) % grid for dy in (0, judge_stride) for dx in (0, judge_stride)]
        if any(jbad[c] for c in corners):
            fall.append(i)
    t_fall = 0.0
    if fall:
        _, yf, t_fall = ftlm_twists(lattice, u, v, tp, [tw[i] for i in fall], betas, mu_min, mu_max, mu_count,
                                    samples, steps, seed)
        yf = yf.reshape(len(fall), nb * nm, 2)
        for k, i in enumerate(fall):
            out[i] = yf[k]
            src[i] = 1
    return {"x": xs, "y": out.reshape(-1, 2), "source": np.repeat(src, nb * nm), "twist_source": src,
            "judges": jidx, "judge_err": jerr, "fallback": fall, "t_surrogate": t_sur, "t_judge": t_judge,
            "t_fallback": t_fall}


# ---------------------------------------------------------------------- CLI
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
    s = sub.add_parser("scan")
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
    s.add_argument("--out", required=True)
    a = ap.parse_args(argv)

    def log(m):
        print(m, flush=True)

    if a.cmd == "generate":
        x, y, secs = generate(a.lattice, a.draws, a.twists, mu_count=a.mu_count, samples=a.samples, steps=a.steps,
                              seed=a.seed, log=log)
        np.savez(a.out, x=x, y=y, ftlm_seconds=secs, lattice=a.lattice)
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
    else:
        sur = Surrogate.load(a.model)
        r = scan(sur, a.lattice, a.u, a.v, a.tp, a.grid, [float(b) for b in a.betas.split(",")],
                 judge_stride=a.judge_stride, tol=a.tol, samples=a.samples, steps=a.steps)
        np.savez(a.out, **{k: np.asarray(v) for k, v in r.items()})
        log(json.dumps({"twists": a.grid ** 2, "judges": len(r["judges"]), "fallback": len(r["fallback"]),
                        "surrogate_twists": int((r["twist_source"] == 0).sum()),
                        "t_surrogate": r["t_surrogate"], "t_judge": r["t_judge"], "t_fallback": r["t_fallback"]}))
    return 0


if __name__ == "__main__":
    sys.exit(_main())
