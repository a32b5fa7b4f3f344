"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where it was
built, the reference harness (oracle/_ref/libftlm_ref.so).

Test infrastructure: used by the tests, the smoke check and the CPU-baseline
leg of bench.py as the checker, never by the product.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libftlm_ref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_ulonglong)


class Model(C.Structure):
    _fields_ = [("lx", C.c_int), ("ly", C.c_int), ("tx", C.c_double), ("ty", C.c_double),
                ("tp", C.c_double), ("phix", C.c_double), ("phiy", C.c_double),
                ("u", C.c_double), ("v", C.c_double)]


def model(lx=2, ly=2, tx=1.0, ty=1.0, tp=0.0, phix=1.0, phiy=1.0, u=8.0, v=0.0):
    return Model(lx, ly, tx, ty, tp, phix, phiy, u, v)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _setup(lib, p):
    """Declare the signatures shared by liboracle (prefix orc_) and the
    reference harness (prefix ref_)."""
    f = lambda n: getattr(lib, p + n)
    f("block_new").restype = C.c_void_p
    f("block_new").argtypes = [C.POINTER(Model), C.c_int, C.c_int, C.c_int, C.c_int]
    f("block_free").argtypes = [C.c_void_p]
    for n in ("block_dim", "block_nnz", "block_full_dim", "block_parents"):
        f(n).restype = C.c_longlong
        f(n).argtypes = [C.c_void_p]
    f("block_reps").argtypes = [C.c_void_p, _up, _up, _ip, _dp]
    f("block_csr").argtypes = [C.c_void_p, _ip, _ip, _dp, _dp]
    f("block_apply").argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    f("mix_u64").restype = C.c_ulonglong
    f("mix_u64").argtypes = [C.c_ulonglong]
    f("block_seed").restype = C.c_ulonglong
    f("block_seed").argtypes = [C.c_ulonglong, C.c_int, C.c_int, C.c_int, C.c_int]
    f("accumulate").argtypes = [_dp, _dp, C.c_longlong, C.c_double, C.c_int, C.c_double, _dp, C.c_int,
                                _dp, _dp, _dp, _dp]
    f("lattice").argtypes = [C.c_int, C.c_int, _ip, _ip, _ip]
    f("twist_phases").argtypes = [C.POINTER(Model), _dp]
    f("diag_energy").restype = C.c_double
    f("diag_energy").argtypes = [C.POINTER(Model), C.c_ulonglong, C.c_ulonglong]
    f("hop_sign").restype = C.c_double
    f("hop_sign").argtypes = [C.c_ulonglong, C.c_int, C.c_int]
    f("combos").restype = C.c_longlong
    f("combos").argtypes = [C.c_int, C.c_int, _up]
    f("linspace").argtypes = [C.c_double, C.c_double, C.c_int, _dp]


class Block:
    """One (N_up, N_down, mx, my) momentum block built by a CPU library."""

    def __init__(self, lib, p, m, nup, ndn, mx, my):
        self.lib, self.p = lib, p
        self.h = getattr(lib, p + "block_new")(C.byref(m), nup, ndn, mx, my)
        if not self.h:
            raise RuntimeError("block construction failed")
        self.dim = getattr(lib, p + "block_dim")(self.h)
        self.nnz = getattr(lib, p + "block_nnz")(self.h)
        self.full_dim = getattr(lib, p + "block_full_dim")(self.h)
        self.parents = getattr(lib, p + "block_parents")(self.h)
        self.particles = nup + ndn

    def __del__(self):
        if getattr(self, "h", None):
            getattr(self.lib, self.p + "block_free")(self.h)
            self.h = None

    def reps(self):
        n = self.dim
        up = np.zeros(n, np.uint64)
        dn = np.zeros(n, np.uint64)
        deg = np.zeros(n, np.int32)
        dnf = np.zeros(n, np.float64)
        getattr(self.lib, self.p + "block_reps")(self.h, _ptr(up, _up), _ptr(dn, _up), _ptr(deg, _ip),
                                                 _ptr(dnf, _dp))
        return up, dn, deg, dnf

    def csr(self):
        rp = np.zeros(self.dim + 1, np.int32)
        col = np.zeros(max(self.nnz, 1), np.int32)
        val = np.zeros(max(self.nnz, 1), np.complex128)
        diag = np.zeros(max(self.dim, 1), np.float64)
        getattr(self.lib, self.p + "block_csr")(self.h, _ptr(rp, _ip), _ptr(col, _ip), _ptr(val, _dp),
                                                _ptr(diag, _dp))
        return rp, col[: self.nnz], val[: self.nnz], diag[: self.dim]

    def apply(self, x):
        """x: complex array of shape (dim,) or (dim, R)."""
        x2 = np.ascontiguousarray(x.reshape(self.dim, -1), dtype=np.complex128)
        y = np.zeros_like(x2)
        getattr(self.lib, self.p + "block_apply")(self.h, x2.shape[1], _ptr(x2, _dp), _ptr(y, _dp))
        return y.reshape(x.shape)

    def dense(self):
        """Dense matrix H[row, col] as the block's CSR stores it."""
        rp, col, val, diag = self.csr()
        H = np.zeros((self.dim, self.dim), np.complex128)
        for i in range(self.dim):
            H[i, i] += diag[i]
            for e in range(rp[i], rp[i + 1]):
                H[i, col[e]] += val[e]
        return H


class _Lib:
    def __init__(self, path, prefix):
        self.lib = C.CDLL(path)
        self.p = prefix
        _setup(self.lib, prefix)

    def block(self, m, nup, ndn, mx, my):
        return Block(self.lib, self.p, m, nup, ndn, mx, my)

    def f(self, name):
        return getattr(self.lib, self.p + name)

    def mix_u64(self, z):
        return self.f("mix_u64")(z)

    def block_seed(self, seed, nup, ndn, mx, my):
        return self.f("block_seed")(seed, nup, ndn, mx, my)

    def lattice(self, lx, ly):
        L = lx * ly
        nb = np.zeros(L * 8, np.int32)
        bonds = np.zeros(4 * L, np.int32)
        nbonds = C.c_int(0)
        self.f("lattice")(lx, ly, _ptr(nb, _ip), _ptr(bonds, _ip), C.byref(nbonds))
        return nb.reshape(L, 8), bonds[: 2 * nbonds.value].reshape(-1, 2)

    def twist_phases(self, m):
        out = np.zeros(8, np.complex128)
        self.f("twist_phases")(C.byref(m), _ptr(out, _dp))
        return out

    def diag_energy(self, m, up, dn):
        return self.f("diag_energy")(C.byref(m), up, dn)

    def hop_sign(self, bits, a, b):
        return self.f("hop_sign")(bits, a, b)

    def combos(self, L, n):
        cnt = self.f("combos")(L, n, None)
        out = np.zeros(max(cnt, 1), np.uint64)
        self.f("combos")(L, n, _ptr(out, _up))
        return out[:cnt]

    def linspace(self, a, b, n):
        out = np.zeros(n, np.float64)
        self.f("linspace")(a, b, n, _ptr(out, _dp))
        return out

    def accumulate(self, eig, w, pref, particles, beta, mu, state):
        """state: dict of emin, z, n1, n2 arrays (updated in place)."""
        eig = np.ascontiguousarray(eig, np.float64)
        wp = None if w is None else _ptr(np.ascontiguousarray(w, np.float64), _dp)
        if w is not None:
            w = np.ascontiguousarray(w, np.float64)
            wp = _ptr(w, _dp)
        mu = np.ascontiguousarray(mu, np.float64)
        self.f("accumulate")(_ptr(eig, _dp), wp, len(eig), pref, particles, beta, _ptr(mu, _dp), len(mu),
                             _ptr(state["emin"], _dp), _ptr(state["z"], _dp), _ptr(state["n1"], _dp),
                             _ptr(state["n2"], _dp))


class Oracle(_Lib):
    def __init__(self, path=ORACLE_SO):
        super().__init__(path, "orc_")
        L = self.lib
        L.orc_start_vector.argtypes = [C.c_ulonglong, C.c_int, C.c_longlong, _dp]
        L.orc_lanczos.restype = C.c_int
        L.orc_lanczos.argtypes = [C.c_void_p, _dp, C.c_ulonglong, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_ritz.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp]
        L.orc_jacobi.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_ftlm_block.restype = C.c_longlong
        L.orc_ftlm_block.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_ulonglong, C.c_int, C.c_int, _dp, _dp]
        L.orc_exact_block.argtypes = [C.c_void_p, _dp]
        L.orc_herm_eigvals.argtypes = [C.c_int, _dp, _dp]
        L.orc_shift_state.argtypes = [C.c_ulonglong, C.c_ulonglong, C.c_int, C.c_int, C.c_int, _up, _up, _ip]
        L.orc_main.restype = C.c_int
        L.orc_main.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.c_int]

    def start_vector(self, bseed, sample, n):
        out = np.zeros(n, np.complex128)
        self.lib.orc_start_vector(bseed, sample, n, _ptr(out, _dp))
        return out

    def lanczos(self, blk, steps, bseed=0, sample=0, start=None, order=0):
        al = np.zeros(steps, np.float64)
        be = np.zeros(steps, np.float64)
        sp = None
        if start is not None:
            start = np.ascontiguousarray(start, np.complex128)
            sp = _ptr(start, _dp)
        m = self.lib.orc_lanczos(blk.h, sp, bseed, sample, steps, order, _ptr(al, _dp), _ptr(be, _dp))
        return al[:m], be[: max(m - 1, 0)]

    def ritz(self, alpha, beta, R):
        m = len(alpha)
        alpha = np.ascontiguousarray(alpha, np.float64)
        beta = np.ascontiguousarray(beta if len(beta) else np.zeros(1), np.float64)
        e = np.zeros(m)
        w = np.zeros(m)
        self.lib.orc_ritz(m, _ptr(alpha, _dp), _ptr(beta, _dp), R, _ptr(e, _dp), _ptr(w, _dp))
        return e, w

    def ftlm_block(self, blk, R, steps, bseed, order=0, threads=0):
        n = R * min(steps, blk.dim)
        e = np.zeros(n)
        w = np.zeros(n)
        cnt = self.lib.orc_ftlm_block(blk.h, R, steps, bseed, order, threads, _ptr(e, _dp), _ptr(w, _dp))
        return e[:cnt], w[:cnt]

    def exact_block(self, blk):
        e = np.zeros(blk.dim)
        self.lib.orc_exact_block(blk.h, _ptr(e, _dp))
        return e

    def herm_eigvals(self, A):
        n = A.shape[0]
        a = np.asfortranarray(A, dtype=np.complex128).ravel(order="F").copy()
        w = np.zeros(n)
        self.lib.orc_herm_eigvals(n, _ptr(a, _dp), _ptr(w, _dp))
        return w

    def shift_state(self, up, dn, j, lx, ly):
        a = C.c_ulonglong(0)
        b = C.c_ulonglong(0)
        s = C.c_int(0)
        self.lib.orc_shift_state(up, dn, j, lx, ly, C.byref(a), C.byref(b), C.byref(s))
        return a.value, b.value, s.value

    def main(self, args, ed=False):
        argv = [b"ftlm_oracle"] + [a.encode() for a in args]
        arr = (C.c_char_p * len(argv))(*argv)
        return self.lib.orc_main(len(argv), arr, 1 if ed else 0)


class Reference(_Lib):
    def __init__(self, path=REF_SO):
        super().__init__(path, "ref_")
        L = self.lib
        L.ref_ftlm_block.restype = C.c_longlong
        L.ref_ftlm_block.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_ulonglong, _dp, _dp]
        L.ref_jacobi.argtypes = [C.c_int, _dp, _dp, _dp]
        L.ref_translate.argtypes = [C.c_ulonglong, C.c_ulonglong, C.c_int, C.c_int, C.c_int, C.c_int, _up, _up,
                                    _ip]
        L.ref_write_results.restype = C.c_int
        L.ref_write_results.argtypes = [C.c_char_p, C.c_int, _dp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.ref_checkpoint_append.restype = C.c_int
        L.ref_checkpoint_append.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_ulonglong, C.c_double, C.c_int, C.c_longlong, _dp, C.c_longlong,
                                            _dp]
        L.ref_checkpoint_count.restype = C.c_longlong
        L.ref_checkpoint_count.argtypes = [C.c_char_p]
        L.ref_metadata_text.restype = C.c_int
        L.ref_metadata_text.argtypes = [C.POINTER(Model), C.c_int, C.c_int, C.c_ulonglong, C.c_ulonglong,
                                        C.c_char_p, C.c_int]
        L.ref_parse_args.restype = C.c_int
        L.ref_parse_args.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.c_char_p, C.c_int]

    def ftlm_block(self, blk, R, steps, bseed, threads=0):
        n = R * min(steps, blk.dim)
        e = np.zeros(n)
        w = np.zeros(n)
        cnt = self.lib.ref_ftlm_block(blk.h, R, steps, threads, bseed, _ptr(e, _dp), _ptr(w, _dp))
        return e[:cnt], w[:cnt]

    def jacobi(self, A):
        n = A.shape[0]
        a = np.ascontiguousarray(A, np.float64)
        e = np.zeros(n)
        v = np.zeros(n)
        self.lib.ref_jacobi(n, _ptr(a, _dp), _ptr(e, _dp), _ptr(v, _dp))
        return e, v

    def translate(self, up, dn, dx, dy, lx, ly):
        a = C.c_ulonglong(0)
        b = C.c_ulonglong(0)
        s = C.c_int(0)
        self.lib.ref_translate(up, dn, dx, dy, lx, ly, C.byref(a), C.byref(b), C.byref(s))
        return a.value, b.value, s.value

    def metadata_text(self, m, samples, steps, threshold, seed):
        buf = C.create_string_buffer(4096)
        n = self.lib.ref_metadata_text(C.byref(m), samples, steps, threshold, seed, buf, 4096)
        return buf.value[:n].decode()

    def parse_args(self, args):
        argv = [b"ftlm_n_vs_mu"] + [a.encode() for a in args]
        arr = (C.c_char_p * len(argv))(*argv)
        err = C.create_string_buffer(1024)
        rc = self.lib.ref_parse_args(len(argv), arr, err, 1024)
        return rc, err.value.decode()


def oracle():
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
    return Oracle()


def reference():
    """The reference harness, or None where it was not built."""
    return Reference() if os.path.exists(REF_SO) else None
