"""MI355X FTLM engine for the extended Hubbard model -- Python binding.

A thin ctypes layer over libftlm_hip.so (the C ABI of include/ftlm.h). The
engine runs on the GPU only: loading fails loudly when the library is not
built, and creating an Engine fails when no HIP device is usable. There is no
CPU fallback.

The Engine methods mirror the reference's hot-path seams (src/main.cpp):
build_sector_basis (622-721) -> sector, build_momentum_block (723-805) ->
block, apply_hamiltonian (833-845) -> apply, run_ftlm_block (996-1113) ->
ftlm_block, diagonalize_block_exact (847-911) -> exact_blocks.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libftlm_hip.so")
HOST_LIB_PATH = os.path.join(PKG_DIR, "libftlm_host.so")
FTLM_EXE = os.path.join(PKG_DIR, "ftlm_n_vs_mu")
ED_EXE = os.path.join(PKG_DIR, "ftlm_ed_n_vs_mu")

REDUCE_SERIAL = 0
REDUCE_TREE = 1
START_HOST = 0
START_DEVICE = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)
_up = C.POINTER(C.c_uint64)
_fp = C.POINTER(C.c_float)


class FtlmModel(C.Structure):
    _fields_ = [("lx", C.c_int32), ("ly", C.c_int32), ("tx", C.c_double), ("ty", C.c_double),
                ("tp", C.c_double), ("phix", C.c_double), ("phiy", C.c_double),
                ("u", C.c_double), ("v", C.c_double)]


class FtlmMlpConfig(C.Structure):
    _fields_ = [("batch", C.c_int32), ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("eps", C.c_float), ("y_mean", C.c_double * 2), ("y_scale", C.c_double * 2)]


class FtlmStats(C.Structure):
    _fields_ = [("ms_apply", C.c_double), ("ms_total", C.c_double), ("n_apply", C.c_int64),
                ("bytes_apply", C.c_double)]


# every symbol declared in include/ftlm.h, with its ctypes signature
SIGNATURES = {
    "ftlm_last_error": (C.c_char_p, []),
    "ftlm_block_seed": (C.c_uint64, [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int]),
    "ftlm_create": (C.c_int, [C.POINTER(FtlmModel), C.c_int, C.POINTER(C.c_void_p)]),
    "ftlm_destroy": (None, [C.c_void_p]),
    "ftlm_sector": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _up, _lp]),
    "ftlm_block": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _lp]),
    "ftlm_block_reps": (C.c_int, [C.c_void_p, _up, _up, _ip, _dp]),
    "ftlm_block_csr": (C.c_int, [C.c_void_p, C.c_int, _lp, _ip, _ip, _dp, _dp]),
    "ftlm_apply": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _dp]),
    "ftlm_start_vector": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, _dp]),
    "ftlm_lanczos": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, C.c_uint64, C.c_int, C.c_int, _dp, _dp, _ip]),
    "ftlm_ritz": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _ip, C.c_int, C.c_int, _dp, _dp]),
    "ftlm_ftlm_block": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_int, _dp, _dp, _lp]),
    "ftlm_prepare": (C.c_int, [C.c_void_p, C.c_int, C.c_uint64]),
    "ftlm_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "ftlm_finish": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _lp]),
    "ftlm_last_stats": (C.c_int, [C.c_void_p, C.POINTER(FtlmStats)]),
    "ftlm_exact_blocks": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _dp, _lp]),
    "ftlm_set_apply_config": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "ftlm_get_apply_config": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip]),
    "ftlm_autotune": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _dp, _ip, _ip, _ip, _dp]),
    "ftlm_set_vector_placement": (C.c_int, [C.c_void_p, C.c_int, C.c_uint64]),
    "ftlm_vector_placement_report": (C.c_int, [C.c_void_p, _ip, _dp, _ip]),
    "ftlm_set_twist": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "ftlm_pack_begin": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_uint64]),
    "ftlm_pack_add": (C.c_int, [C.c_void_p, C.c_uint64, _lp]),
    "ftlm_pack_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "ftlm_pack_result": (C.c_int, [C.c_void_p, C.c_int64, _dp, _dp, _lp]),
    "ftlm_pack_tune": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _dp, _ip, _ip, _ip, _dp]),
    "ftlm_pack_info": (C.c_int, [C.c_void_p, _lp, _up]),
    "ftlm_thermo": (C.c_int, [C.c_void_p, C.c_int64, _lp, _dp, _dp, _ip, _dp, C.c_int, _dp, C.c_int, _dp, _dp]),
    "ftlm_thermo_noalloc": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                      C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ftlm_memprobe": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.c_int, C.c_int, C.c_int, _dp, _dp]),
    "ftlm_set_start_source": (C.c_int, [C.c_void_p, C.c_int]),
    "ftlm_exact_queue_begin": (C.c_int, [C.c_void_p, C.c_uint64]),
    "ftlm_exact_queue_add": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _lp]),
    "ftlm_exact_queue_run": (C.c_int, [C.c_void_p]),
    "ftlm_exact_queue_result": (C.c_int, [C.c_void_p, C.c_int64, _dp, _lp]),
    # learned surrogate
    "ftlm_mlp_param_count": (C.c_int64, []),
    "ftlm_mlp_layout": (C.c_int, [_lp, _lp, _ip, _ip]),
    "ftlm_mlp_encode": (C.c_int, [C.c_int64, _dp, _fp]),
    "ftlm_mlp_create": (C.c_int, [C.POINTER(FtlmMlpConfig), C.c_int, C.POINTER(C.c_void_p)]),
    "ftlm_mlp_destroy": (None, [C.c_void_p]),
    "ftlm_mlp_set_params": (C.c_int, [C.c_void_p, _fp]),
    "ftlm_mlp_get_params": (C.c_int, [C.c_void_p, _fp]),
    "ftlm_mlp_set_lr": (C.c_int, [C.c_void_p, C.c_float]),
    "ftlm_mlp_set_data": (C.c_int, [C.c_void_p, C.c_int64, _dp, _dp]),
    "ftlm_mlp_train": (C.c_int, [C.c_void_p, C.c_int, C.c_uint64, _fp]),
    "ftlm_mlp_gradients": (C.c_int, [C.c_void_p, C.c_int64, _dp, _dp, _fp, _fp]),
    "ftlm_mlp_step": (C.c_int, [C.c_void_p, C.c_int64, _dp, _dp, _fp]),
    "ftlm_mlp_predict": (C.c_int, [C.c_void_p, C.c_int64, _dp, _dp]),
    "ftlm_mlp_gemm": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _fp, _fp, _fp]),
}

PACK_FULL = 1

_lib = None


def load_library(path=LIB_PATH):
    """Load libftlm_hip.so (raises if it is not built: no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run `make -C {PKG_DIR}` (or __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a, t):
    return a.ctypes.data_as(t)


class FtlmError(RuntimeError):
    pass


def block_seed(seed, nup, ndn, mx, my):
    """make_block_seed (main.cpp:520-533)."""
    return load_library().ftlm_block_seed(seed, nup, ndn, mx, my)


class Engine:
    """One GPU context: a lattice model, a current sector and block."""

    def __init__(self, lx=2, ly=2, tx=1.0, ty=1.0, tp=0.0, phix=1.0, phiy=1.0, u=8.0, v=0.0, device=0):
        self.lib = load_library()
        self.model = FtlmModel(lx, ly, tx, ty, tp, phix, phiy, u, v)
        self.L = lx * ly
        h = C.c_void_p()
        self._check(self.lib.ftlm_create(C.byref(self.model), device, C.byref(h)))
        self.h = h
        self.nup = self.ndn = None
        self.dim = 0

    def _check(self, rc):
        if rc != 0:
            raise FtlmError(f"ftlm error {rc}: {self.lib.ftlm_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.ftlm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sector(self, nup, ndn):
        fd = C.c_uint64(0)
        par = C.c_int64(0)
        self._check(self.lib.ftlm_sector(self.h, nup, ndn, C.byref(fd), C.byref(par)))
        self.nup, self.ndn = nup, ndn
        return fd.value, par.value

    def block(self, mx, my):
        d = C.c_int64(0)
        self._check(self.lib.ftlm_block(self.h, mx, my, C.byref(d)))
        self.dim = d.value
        return self.dim

    def reps(self):
        n = self.dim
        up = np.zeros(n, np.uint64)
        dn = np.zeros(n, np.uint64)
        deg = np.zeros(n, np.int32)
        dnf = np.zeros(n, np.float64)
        self._check(self.lib.ftlm_block_reps(self.h, _p(up, _up), _p(dn, _up), _p(deg, _ip), _p(dnf, _dp)))
        return up, dn, deg, dnf

    def csr(self, all_dirs=True):
        nnz = C.c_int64(0)
        self._check(self.lib.ftlm_block_csr(self.h, int(all_dirs), C.byref(nnz), None, None, None, None))
        n = self.dim
        rp = np.zeros(n + 1, np.int32)
        col = np.zeros(max(nnz.value, 1), np.int32)
        val = np.zeros(max(nnz.value, 1), np.complex128)
        diag = np.zeros(max(n, 1), np.float64)
        self._check(self.lib.ftlm_block_csr(self.h, int(all_dirs), C.byref(nnz), _p(rp, _ip), _p(col, _ip),
                                            _p(val, _dp), _p(diag, _dp)))
        return rp, col[: nnz.value], val[: nnz.value], diag[:n]

    def apply(self, x, kernel=0):
        """x: complex (dim,) or (dim, R) -> H x (same shape); kernel 0 uses
        the apply configuration, 1 the on-the-fly one-row kernel (R = 1)."""
        x2 = np.ascontiguousarray(np.asarray(x).reshape(self.dim, -1), dtype=np.complex128)
        y = np.zeros_like(x2)
        self._check(self.lib.ftlm_apply(self.h, x2.shape[1], kernel, _p(x2, _dp), _p(y, _dp)))
        return y.reshape(np.shape(x))

    def start_vector(self, bseed, sample):
        out = np.zeros(self.dim, np.complex128)
        self._check(self.lib.ftlm_start_vector(self.h, bseed, sample, _p(out, _dp)))
        return out

    def lanczos(self, R, steps, start=None, bseed=0, reduce=REDUCE_TREE, per_vector=False):
        """alpha [steps][R], beta [steps][R], m [R] of R vectors."""
        a = np.zeros((steps, R))
        b = np.zeros((steps, R))
        m = np.zeros(R, np.int32)
        sp = None
        if start is not None:
            start = np.ascontiguousarray(np.asarray(start).reshape(self.dim, R), dtype=np.complex128)
            sp = _p(start, _dp)
        self._check(self.lib.ftlm_lanczos(self.h, R, steps, sp, bseed, reduce, int(per_vector), _p(a, _dp),
                                          _p(b, _dp), _p(m, _ip)))
        return a, b, m

    def ritz(self, alpha, beta, m, wdiv, method=0):
        alpha = np.ascontiguousarray(alpha, np.float64)
        beta = np.ascontiguousarray(beta, np.float64)
        m = np.ascontiguousarray(m, np.int32)
        steps, R = alpha.shape
        e = np.zeros((R, steps))
        w = np.zeros((R, steps))
        self._check(self.lib.ftlm_ritz(self.h, R, steps, _p(alpha, _dp), _p(beta, _dp), _p(m, _ip), wdiv, method,
                                       _p(e, _dp), _p(w, _dp)))
        return e, w

    def ftlm_block(self, R, steps, bseed, reduce=REDUCE_TREE, per_vector=False):
        n = R * min(steps, self.dim)
        e = np.zeros(max(n, 1))
        w = np.zeros(max(n, 1))
        cnt = C.c_int64(0)
        self._check(self.lib.ftlm_ftlm_block(self.h, R, steps, bseed, reduce, int(per_vector), _p(e, _dp),
                                             _p(w, _dp), C.byref(cnt)))
        return e[: cnt.value], w[: cnt.value]

    def last_stats(self):
        """Device-event timings of the last Lanczos run (ftlm_run or the
        batched runs of ftlm_block)."""
        s = FtlmStats()
        self._check(self.lib.ftlm_last_stats(self.h, C.byref(s)))
        return {"ms_apply": s.ms_apply, "ms_total": s.ms_total, "n_apply": s.n_apply,
                "bytes_apply": s.bytes_apply}

    def prepare(self, R, bseed):
        self._check(self.lib.ftlm_prepare(self.h, R, bseed))

    def run(self, R, steps, reduce=REDUCE_TREE, per_vector_column=-1):
        self._check(self.lib.ftlm_run(self.h, R, steps, reduce, per_vector_column))
        s = FtlmStats()
        self._check(self.lib.ftlm_last_stats(self.h, C.byref(s)))
        return {"ms_apply": s.ms_apply, "ms_total": s.ms_total, "n_apply": s.n_apply,
                "bytes_apply": s.bytes_apply}

    def finish(self, R, steps):
        n = R * min(steps, self.dim)
        e = np.zeros(max(n, 1))
        w = np.zeros(max(n, 1))
        cnt = C.c_int64(0)
        self._check(self.lib.ftlm_finish(self.h, R, steps, _p(e, _dp), _p(w, _dp), C.byref(cnt)))
        return e[: cnt.value], w[: cnt.value]

    def set_apply_config(self, path=0, W=-1):
        """path: 0 automatic, 1 hop list, 2 on the fly; W: -1 automatic, 0
        untiled, >0 column tile width."""
        self._check(self.lib.ftlm_set_apply_config(self.h, path, W))

    def apply_config(self, R):
        p = C.c_int32(0)
        w = C.c_int32(0)
        self._check(self.lib.ftlm_get_apply_config(self.h, R, C.byref(p), C.byref(w)))
        return p.value, w.value

    def set_vector_placement(self, candidates=8, min_bytes=256 << 20):
        """Placement tuning of the Lanczos vectors (0 candidates: off)."""
        self._check(self.lib.ftlm_set_vector_placement(self.h, candidates, min_bytes))

    def vector_placement_report(self):
        """(apply ms per candidate of the last tuning, chosen candidates)."""
        n = C.c_int32(16)
        ms = np.zeros(16)
        pick = np.zeros(3, np.int32)
        self._check(self.lib.ftlm_vector_placement_report(self.h, C.byref(n), _p(ms, _dp), _p(pick, _ip)))
        return ms[:n.value].copy(), (pick.tolist() if n.value else [])

    def autotune(self, R):
        """Time the apply candidates on the current block; returns
        ((path, W), best ms, {"path/W": ms})."""
        bp = C.c_int32(0)
        bw = C.c_int32(0)
        bms = C.c_double(0)
        n = C.c_int32(16)
        cp = np.zeros(16, np.int32)
        cw = np.zeros(16, np.int32)
        cm = np.zeros(16)
        self._check(self.lib.ftlm_autotune(self.h, R, C.byref(bp), C.byref(bw), C.byref(bms), C.byref(n),
                                           _p(cp, _ip), _p(cw, _ip), _p(cm, _dp)))
        names = {1: "hoplist", 2: "onthefly"}
        return ((bp.value, bw.value), bms.value,
                {f"{names[int(cp[i])]}/W{int(cw[i])}": round(float(cm[i]), 3) for i in range(n.value)})

    def set_twist(self, phix, phiy):
        """Change the twist; the current block must be selected again."""
        self._check(self.lib.ftlm_set_twist(self.h, phix, phiy))
        self.model.phix = phix
        self.model.phiy = phiy

    # ---- packed blocks (one persistent launch over many blocks) ----
    def pack_begin(self, R, steps, budget_bytes=0):
        self._check(self.lib.ftlm_pack_begin(self.h, R, steps, budget_bytes))
        self._pack = {"R": R, "steps": steps, "dims": []}

    def pack_add(self, bseed):
        """Add the current block; returns its index, or None when the pack is
        full (nothing added)."""
        idx = C.c_int64(-1)
        rc = self.lib.ftlm_pack_add(self.h, bseed, C.byref(idx))
        if rc == PACK_FULL:
            return None
        self._check(rc)
        self._pack["dims"].append(self.dim)
        return idx.value

    def pack_run(self, width=0, wg_per_cu=0, method=0):
        """method 0 Jacobi Ritz (bit-identical to the reference), 1 QL.
        Returns the device-event stats of the run."""
        self._check(self.lib.ftlm_pack_run(self.h, width, wg_per_cu, method))
        s = FtlmStats()
        self._check(self.lib.ftlm_last_stats(self.h, C.byref(s)))
        return {"ms_total": s.ms_total, "n_apply": s.n_apply, "bytes_apply": s.bytes_apply}

    def pack_result(self, index):
        dims = self._pack["dims"]
        d = dims[index] if 0 <= index < len(dims) else 1
        n = self._pack["R"] * min(self._pack["steps"], d)
        e = np.zeros(max(n, 1))
        w = np.zeros(max(n, 1))
        cnt = C.c_int64(0)
        self._check(self.lib.ftlm_pack_result(self.h, index, _p(e, _dp), _p(w, _dp), C.byref(cnt)))
        return e[: cnt.value], w[: cnt.value]

    def pack_tune(self, tune_steps=4):
        """Time (width, workgroups per CU) candidates; returns ((W, wg), ms, {"W/wg": ms})."""
        bw = C.c_int32(0)
        bg = C.c_int32(0)
        bms = C.c_double(0)
        n = C.c_int32(32)
        cw = np.zeros(32, np.int32)
        cg = np.zeros(32, np.int32)
        cm = np.zeros(32)
        self._check(self.lib.ftlm_pack_tune(self.h, tune_steps, C.byref(bw), C.byref(bg), C.byref(bms), C.byref(n),
                                            _p(cw, _ip), _p(cg, _ip), _p(cm, _dp)))
        return ((bw.value, bg.value), bms.value,
                {f"W{int(cw[i])}/wg{int(cg[i])}": round(float(cm[i]), 3) for i in range(n.value)})

    def pack_info(self):
        n = C.c_int64(0)
        b = C.c_uint64(0)
        self._check(self.lib.ftlm_pack_info(self.h, C.byref(n), C.byref(b)))
        return n.value, b.value

    def memprobe(self, kind, rows, width, k=36, reps=3):
        """HBM probe with the batched apply's shapes (kind 0 read, 1 copy, 2
        random gather of k rows per output row): (ms per launch, bytes)."""
        ms = C.c_double(0)
        by = C.c_double(0)
        self._check(self.lib.ftlm_memprobe(self.h, kind, rows, width, k, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def set_start_source(self, source):
        """START_HOST (libstdc++ stream on host threads, bit-identical to the
        reference) or START_DEVICE (the same stream drawn on the GPU)."""
        self._check(self.lib.ftlm_set_start_source(self.h, source))

    # ---- queued exact blocks (one eigensolver launch for many sectors/twists) ----
    def exact_queue_begin(self, budget_bytes=0):
        self._check(self.lib.ftlm_exact_queue_begin(self.h, budget_bytes))
        self._exq = []

    def exact_queue_add(self, mx, my):
        """Queue block (mx, my) of the current sector and twist; returns its
        index, or None when the queue is full (nothing added)."""
        idx = C.c_int64(-1)
        rc = self.lib.ftlm_exact_queue_add(self.h, mx, my, C.byref(idx))
        if rc == PACK_FULL:
            return None
        self._check(rc)
        return idx.value

    def exact_queue_run(self):
        self._check(self.lib.ftlm_exact_queue_run(self.h))

    def exact_queue_result(self, index):
        cnt = C.c_int64(0)
        self._check(self.lib.ftlm_exact_queue_result(self.h, index, None, C.byref(cnt)))
        e = np.zeros(max(cnt.value, 1))
        self._check(self.lib.ftlm_exact_queue_result(self.h, index, _p(e, _dp), C.byref(cnt)))
        return e[: cnt.value]

    def thermo(self, spectra, betas, mu):
        """Device log-sum-exp sums (ftlm_thermo) of spectra = [(eig, w or
        None for an exact block, particles, prefactor)] for every beta and mu.
        Returns arrays (emin, z, n1, n2), each [len(betas)][len(mu)]."""
        off = np.zeros(len(spectra) + 1, np.int64)
        for i, (e, _, _, _) in enumerate(spectra):
            off[i + 1] = off[i] + len(e)
        eig = np.ascontiguousarray(np.concatenate([np.asarray(s[0], np.float64) for s in spectra])
                                   if spectra else np.zeros(0))
        w = np.ascontiguousarray(np.concatenate([np.ones(len(s[0])) if s[1] is None else np.asarray(s[1], np.float64)
                                                 for s in spectra]) if spectra else np.zeros(0))
        part = np.array([s[2] for s in spectra], np.int32)
        pref = np.array([s[3] for s in spectra], np.float64)
        betas = np.ascontiguousarray(betas, np.float64)
        mu = np.ascontiguousarray(mu, np.float64)
        out = np.zeros((betas.size, mu.size, 4))
        self._check(self.lib.ftlm_thermo(self.h, len(spectra), _p(off, _lp), _p(eig, _dp), _p(w, _dp),
                                         _p(part, _ip), _p(pref, _dp), betas.size, _p(betas, _dp), mu.size,
                                         _p(mu, _dp), _p(out, _dp)))
        return out[..., 0], out[..., 1], out[..., 2], out[..., 3]

    def exact_blocks(self, ks):
        """ks: list of (mx, my) of the current sector -> list of eigenvalue arrays."""
        mx = np.array([k[0] for k in ks], np.int32)
        my = np.array([k[1] for k in ks], np.int32)
        dims = []
        for a, b in ks:
            dims.append(self.block(int(a), int(b)))
        tot = int(sum(dims))
        eig = np.zeros(max(tot, 1))
        off = np.zeros(len(ks) + 1, np.int64)
        self._check(self.lib.ftlm_exact_blocks(self.h, len(ks), _p(mx, _ip), _p(my, _ip), _p(eig, _dp),
                                               _p(off, _lp)))
        return [eig[off[i]:off[i + 1]].copy() for i in range(len(ks))]
