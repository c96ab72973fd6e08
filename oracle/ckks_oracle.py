"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product path never does. Every function takes and returns
numpy uint32 arrays in the reference's limb-major layout [limb][coeff].
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_ip = ctypes.POINTER(ctypes.c_int)

OR_MODDOWN, OR_RESCALE, OR_MODDOWN_RESCALE = 0, 1, 2


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB_PATH):
        raise FileNotFoundError(f"{_LIB_PATH} missing: run make -C oracle")
    lib = ctypes.CDLL(_LIB_PATH)
    sig = {
        "or_mulmod": (ctypes.c_uint32, [ctypes.c_uint32] * 3),
        "or_powmod": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_int64, ctypes.c_uint32]),
        "or_invmod": (ctypes.c_uint32, [ctypes.c_uint32] * 2),
        "or_find_generator": (ctypes.c_uint32, [ctypes.c_uint32]),
        "or_primitive_root": (ctypes.c_uint32, [ctypes.c_int64, ctypes.c_uint32]),
        "or_inv_mod_base": (ctypes.c_int32, [ctypes.c_uint32]),
        "or_to_mont": (ctypes.c_uint32, [ctypes.c_uint32] * 2),
        "or_mont_reduce_lazy": (ctypes.c_int32, [ctypes.c_int64, ctypes.c_uint32, ctypes.c_int32]),
        "or_mont_reduce": (ctypes.c_uint32, [ctypes.c_int64, ctypes.c_uint32, ctypes.c_int32]),
        "or_mult_mont": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32]),
        "or_normalize": (ctypes.c_int32, [ctypes.c_uint32] * 2),
        "or_bitrev": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_int]),
        "or_param_create": (ctypes.c_void_p, [ctypes.c_int, _u32p, ctypes.c_int, _u32p, ctypes.c_int,
                                               _u32p, ctypes.c_int, _ip, ctypes.c_int]),
        "or_param_destroy": (None, [ctypes.c_void_p]),
        "or_param_info": (None, [ctypes.c_void_p, _ip]),
        "or_prime": (ctypes.c_uint32, [ctypes.c_void_p, ctypes.c_int]),
        "or_prime_qinv": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int]),
        "or_prime_psi": (ctypes.c_uint32, [ctypes.c_void_p, ctypes.c_int]),
        "or_level_np": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip, _ip]),
        "or_np_level": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
        "or_np_prime": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "or_galois_factor": (ctypes.c_uint32, [ctypes.c_void_p, ctypes.c_int]),
        "or_psi_rev_mont": (_u32p, [ctypes.c_void_p, ctypes.c_int]),
        "or_psi_inv_rev_mont": (_u32p, [ctypes.c_void_p, ctypes.c_int]),
        "or_p_prod_mont": (ctypes.c_uint32, [ctypes.c_void_p, ctypes.c_int]),
        "or_set_num_threads": (None, [ctypes.c_int]),
        "or_get_num_threads": (ctypes.c_int, []),
        "or_fill_uniform": (None, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_uint64, _u32p]),
        "or_fill_uniform_primes": (None, [ctypes.c_void_p, _ip, ctypes.c_int, ctypes.c_uint64, _u32p]),
        "or_ntt_limb": (None, [ctypes.c_void_p, ctypes.c_int, _u32p, ctypes.c_int]),
        "or_intt_raw_limb": (None, [ctypes.c_void_p, ctypes.c_int, _u32p]),
        "or_ntt": (None, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u32p, ctypes.c_int]),
        "or_intt": (None, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u32p, ctypes.c_int]),
        "or_intt_mult_const": (None, [ctypes.c_void_p, _ip, ctypes.c_int, _u32p, _u32p, ctypes.c_int, _i32p]),
        "or_bconv_tables": (None, [ctypes.c_void_p, _ip, ctypes.c_int, _ip, ctypes.c_int, _ip, ctypes.c_int,
                                   _u32p, _i32p]),
        "or_bconv": (None, [ctypes.c_void_p, _ip, ctypes.c_int, _i32p, _ip, ctypes.c_int, _i32p, _u32p]),
        "or_modup_beta": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
        "or_modup": (None, [ctypes.c_void_p, ctypes.c_int, _u32p, _u32p]),
        "or_moddown": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _u32p, _u32p]),
        "or_add": (None, [ctypes.c_void_p, _ip, ctypes.c_int, _u32p, _u32p, _u32p]),
        "or_sub": (None, [ctypes.c_void_p, _ip, ctypes.c_int, _u32p, _u32p, _u32p]),
        "or_neg": (None, [ctypes.c_void_p, _ip, ctypes.c_int, _u32p, _u32p]),
        "or_mult": (None, [ctypes.c_void_p, _ip, ctypes.c_int, _u32p, _u32p, _u32p]),
        "or_mult_const": (None, [ctypes.c_void_p, _ip, ctypes.c_int, _u32p, _u32p, _u32p]),
        "or_tensor": (None, [ctypes.c_void_p, _ip, ctypes.c_int] + [_u32p] * 7),
        "or_permute": (None, [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, _u32p, _u32p]),
        "or_mult_key_no_moddown": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [_u32p] * 7),
        "or_hmult": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [_u32p] * 6 + [ctypes.c_int, _u32p, _u32p]),
        "or_hrot": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [_u32p] * 4 + [ctypes.c_int, _u32p, _u32p]),
        "or_hconj": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [_u32p] * 6),
        "or_mult_key_short_base": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
                                   + [_u32p] * 6),
        "or_mod_up_to_max": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int] + [_u32p] * 8),
        "or_hoisted_baby_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [_u32p] * 4 +
                                 [ctypes.c_int, _u32p, _u32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _p(a, ty=_u32p):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ty)


def _ints(lst):
    arr = (ctypes.c_int * max(1, len(lst)))(*lst)
    return arr


def set_threads(n):
    lib.or_set_num_threads(int(n))


class Param:
    """Oracle parameter set: prime chain, level table and per-prime tables."""

    def __init__(self, log_n, main_primes, aux_primes, ter_primes=(), level_config=None):
        if any(int(q) >= 2 ** 31 for q in list(main_primes) + list(aux_primes) + list(ter_primes)):
            raise ValueError("primes must be below 2^31 (32-bit words)")
        main = np.ascontiguousarray(main_primes, dtype=np.uint32)
        ter = np.ascontiguousarray(ter_primes, dtype=np.uint32)
        aux = np.ascontiguousarray(aux_primes, dtype=np.uint32)
        if level_config is None:
            level_config = [[i + 1, 0] for i in range(len(main))]
        cfg = np.ascontiguousarray(np.array(level_config, dtype=np.int32).reshape(-1))
        self._h = lib.or_param_create(int(log_n), _p(main), len(main),
                                      _p(ter) if len(ter) else None, len(ter),
                                      _p(aux), len(aux),
                                      cfg.ctypes.data_as(_ip), len(level_config))
        if not self._h:
            raise ValueError("invalid parameter set")
        info = (ctypes.c_int * 10)()
        lib.or_param_info(self._h, info)
        (self.log_n, self.n, self.max_main, self.max_ter, self.alpha, self.L,
         self.dnum, self.max_level, self.num_primes, _) = list(info)
        self.primes = np.array([lib.or_prime(self._h, i) for i in range(self.num_primes)],
                               dtype=np.uint64)

    @classmethod
    def from_json(cls, d):
        return cls(d["log_degree"], d["main_primes"], d["auxiliary_primes"],
                   d.get("terminal_primes", []), d["level_config"])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.or_param_destroy(h)
            self._h = None

    # ---- level helpers ------------------------------------------------
    def level_np(self, level):
        m, t = ctypes.c_int(), ctypes.c_int()
        if lib.or_level_np(self._h, level, ctypes.byref(m), ctypes.byref(t)) != 0:
            raise ValueError("bad level")
        return m.value, t.value

    def num_q(self, level):
        m, t = self.level_np(level)
        return m + t

    def np_primes(self, num_main, num_ter, num_aux=0):
        return [lib.or_np_prime(self._h, num_main, num_ter, t)
                for t in range(num_main + num_ter + num_aux)]

    def level_primes(self, level, with_aux=False):
        m, t = self.level_np(level)
        return self.np_primes(m, t, self.alpha if with_aux else 0)

    def prime(self, i):
        return int(self.primes[i])

    def qinv(self, i):
        return lib.or_prime_qinv(self._h, i)

    def psi(self, i):
        return lib.or_prime_psi(self._h, i)

    def galois(self, r):
        return lib.or_galois_factor(self._h, r)

    def p_prod_mont(self, i):
        return lib.or_p_prod_mont(self._h, i)

    def modup_beta(self, level):
        return lib.or_modup_beta(self._h, level)

    def psi_rev_mont(self, i):
        ptr = lib.or_psi_rev_mont(self._h, i)
        return np.ctypeslib.as_array(ptr, shape=(self.n,)).copy()

    def psi_inv_rev_mont(self, i):
        ptr = lib.or_psi_inv_rev_mont(self._h, i)
        return np.ctypeslib.as_array(ptr, shape=(self.n,)).copy()

    # ---- data ---------------------------------------------------------
    def uniform(self, pidx, seed):
        out = np.empty((len(pidx), self.n), dtype=np.uint32)
        lib.or_fill_uniform_primes(self._h, _ints(pidx), len(pidx), ctypes.c_uint64(seed), _p(out))
        return out

    # ---- transforms ---------------------------------------------------
    def ntt(self, data, pidx, mont_conv=False):
        out = np.ascontiguousarray(data, dtype=np.uint32).copy()
        for t, pi in enumerate(pidx):
            row = out[t]
            lib.or_ntt_limb(self._h, pi, _p(row), int(mont_conv))
        return out

    def ntt_np(self, data, num_main, num_ter, num_aux=0, mont_conv=False):
        out = np.ascontiguousarray(data, dtype=np.uint32).copy()
        lib.or_ntt(self._h, num_main, num_ter, num_aux, _p(out), int(mont_conv))
        return out

    def intt_np(self, data, num_main, num_ter, num_aux=0, mont_conv=True):
        out = np.ascontiguousarray(data, dtype=np.uint32).copy()
        lib.or_intt(self._h, num_main, num_ter, num_aux, _p(out), int(mont_conv))
        return out

    def intt_raw(self, data, pidx):
        out = np.ascontiguousarray(data, dtype=np.uint32).copy()
        for t, pi in enumerate(pidx):
            lib.or_intt_raw_limb(self._h, pi, _p(out[t]))
        return out

    def intt_mult_const(self, data, pidx, consts, normalize):
        src = np.ascontiguousarray(data, dtype=np.uint32)
        c = np.ascontiguousarray(consts, dtype=np.uint32)
        out = np.empty(src.shape, dtype=np.int32)
        lib.or_intt_mult_const(self._h, _ints(pidx), len(pidx), _p(src), _p(c), int(normalize),
                               out.ctypes.data_as(_i32p))
        return out

    def bconv_tables(self, src_pidx, dst_pidx, restore_pidx=()):
        s, d = len(src_pidx), len(dst_pidx)
        c1 = np.empty(s, dtype=np.uint32)
        tab = np.empty(d * s, dtype=np.int32)
        lib.or_bconv_tables(self._h, _ints(src_pidx), s, _ints(dst_pidx), d,
                            _ints(restore_pidx), len(restore_pidx), _p(c1),
                            tab.ctypes.data_as(_i32p))
        return c1, tab.reshape(d, s)

    def bconv(self, src, src_pidx, dst_pidx, table):
        src = np.ascontiguousarray(src, dtype=np.int32)
        tab = np.ascontiguousarray(table, dtype=np.int32)
        out = np.empty((len(dst_pidx), self.n), dtype=np.uint32)
        lib.or_bconv(self._h, _ints(src_pidx), len(src_pidx), src.ctypes.data_as(_i32p),
                     _ints(dst_pidx), len(dst_pidx), tab.ctypes.data_as(_i32p), _p(out))
        return out

    def modup(self, src, level):
        src = np.ascontiguousarray(src, dtype=np.uint32)
        beta = self.modup_beta(level)
        tot = self.num_q(level) + self.alpha
        out = np.empty((beta, tot, self.n), dtype=np.uint32)
        lib.or_modup(self._h, level, _p(src), _p(out))
        return out

    def moddown(self, src, level, kind):
        src = np.ascontiguousarray(src, dtype=np.uint32)
        if kind == OR_MODDOWN:
            nq = self.num_q(level)
        else:
            nq = self.num_q(level - 1)
        out = np.empty((nq, self.n), dtype=np.uint32)
        rc = lib.or_moddown(self._h, level, kind, _p(src), _p(out))
        if rc != 0:
            raise ValueError("or_moddown failed")
        return out

    # ---- element-wise -------------------------------------------------
    def _ew(self, fn, pidx, *arrs):
        arrs = [np.ascontiguousarray(a, dtype=np.uint32) for a in arrs]
        out = np.empty(arrs[0].shape, dtype=np.uint32)
        fn(self._h, _ints(pidx), len(pidx), *[_p(a) for a in arrs], _p(out))
        return out

    def add(self, x, y, pidx):
        return self._ew(lib.or_add, pidx, x, y)

    def sub(self, x, y, pidx):
        return self._ew(lib.or_sub, pidx, x, y)

    def neg(self, x, pidx):
        return self._ew(lib.or_neg, pidx, x)

    def mult(self, x, y, pidx):
        return self._ew(lib.or_mult, pidx, x, y)

    def mult_const(self, x, consts, pidx):
        x = np.ascontiguousarray(x, dtype=np.uint32)
        c = np.ascontiguousarray(consts, dtype=np.uint32)
        out = np.empty(x.shape, dtype=np.uint32)
        lib.or_mult_const(self._h, _ints(pidx), len(pidx), _p(x), _p(c), _p(out))
        return out

    def add_const(self, x, consts, pidx, mode=0):
        """AddConst / SubConst / SubOppositeConst (src/core/ElementWise.cu:145-223):
        mode 0: x + c, 1: x - c, 2: c - x, c one value per limb, in [0, q)."""
        q = np.array([self.prime(i) for i in pidx], dtype=np.int64)[:, None]
        xv = np.asarray(x, dtype=np.int64)
        c = np.asarray(consts, dtype=np.int64)[:, None]
        r = (xv + c) if mode == 0 else (xv - c) if mode == 1 else (c - xv)
        return (r % q).astype(np.uint32)

    def mult_imaginary_unit(self, x, pidx):
        """MultImaginaryUnit (src/core/ElementWise.cu:625-658): NTT position k
        times psi^(N/2), negated for k >= N/2 (the NTT form of X^(N/2))."""
        out = np.empty(np.shape(x), dtype=np.uint32)
        half = self.n // 2
        for t, pi in enumerate(pidx):
            q = self.prime(pi)
            i_unit = pow(self.psi(pi), half, q)
            row = np.asarray(x[t], dtype=np.int64)
            out[t, :half] = (row[:half] * i_unit) % q
            out[t, half:] = (row[half:] * (q - i_unit)) % q
        return out

    def modup_to_max(self, src, base_pidx, dst_pidx):
        """ModUpToMax1/2 (src/core/ElementWise.cu:579-623): coefficient-domain
        limbs of a one- or two-prime base lifted to every destination prime.
        Two primes: v = s0 q1 + s1 q0, brought from [0, 2 q0 q1) into
        (-q0 q1 / 2, q0 q1 / 2]; one prime: s0 mod t (not centred)."""
        s = np.asarray(src, dtype=np.int64)
        qd = np.array([self.prime(i) for i in dst_pidx], dtype=np.int64)[:, None]
        if len(base_pidx) == 2:
            q0, q1 = self.prime(base_pidx[0]), self.prime(base_pidx[1])
            qq = q0 * q1  # < 2^62, and v < 2^63: int64 holds every value
            v = s[0] * q1 + s[1] * q0
            v = np.where(v >= qq, v - qq, v)
            v = np.where(v > qq // 2, v - qq, v)
            return (v[None, :] % qd).astype(np.uint32)
        return (s[0][None, :] % qd).astype(np.uint32)

    def tensor(self, b1, a1, b2, a2, pidx):
        ins = [np.ascontiguousarray(a, dtype=np.uint32) for a in (b1, a1, b2, a2)]
        outs = [np.empty(ins[0].shape, dtype=np.uint32) for _ in range(3)]
        lib.or_tensor(self._h, _ints(pidx), len(pidx), *[_p(a) for a in ins], *[_p(o) for o in outs])
        return outs

    def permute(self, x, galois_factor):
        x = np.ascontiguousarray(x, dtype=np.uint32)
        out = np.empty(x.shape, dtype=np.uint32)
        lib.or_permute(self._h, x.shape[0], ctypes.c_uint32(galois_factor), _p(x), _p(out))
        return out

    # ---- key switching ------------------------------------------------
    def mult_key_no_moddown(self, level, x, evk_b, evk_a, add_b=None, add_a=None):
        tot = self.num_q(level) + self.alpha
        ab = np.empty((tot, self.n), dtype=np.uint32)
        aa = np.empty((tot, self.n), dtype=np.uint32)
        args = [np.ascontiguousarray(v, dtype=np.uint32) if v is not None else None
                for v in (x, add_b, add_a, evk_b, evk_a)]
        rc = lib.or_mult_key_no_moddown(self._h, level, *[_p(a) for a in args], _p(ab), _p(aa))
        if rc != 0:
            raise ValueError("or_mult_key_no_moddown failed")
        return ab, aa

    def hmult(self, level, ct1, ct2, evk, rescale=True):
        nq = self.num_q(level - 1 if rescale else level)
        ob = np.empty((nq, self.n), dtype=np.uint32)
        oa = np.empty((nq, self.n), dtype=np.uint32)
        ins = [np.ascontiguousarray(v, dtype=np.uint32) for v in (ct1[0], ct1[1], ct2[0], ct2[1],
                                                                   evk[0], evk[1])]
        rc = lib.or_hmult(self._h, level, *[_p(a) for a in ins], int(rescale), _p(ob), _p(oa))
        if rc != 0:
            raise ValueError("or_hmult failed")
        return ob, oa

    def mult_key_short_base(self, num_main, num_ter, ct, key):
        """MultKey at level -1 (sparse-secret encapsulation's dense -> sparse
        switch, src/core/Context.cpp:567-607 with the level -1
        ModSwitchHandler, src/core/ModSwitch.cu:175-232): ct over the short
        base's num_q limbs, key = (b, a) over 2 num_q limbs"""
        nq = num_main + num_ter
        ob = np.empty((nq, self.n), dtype=np.uint32)
        oa = np.empty((nq, self.n), dtype=np.uint32)
        ins = [np.ascontiguousarray(v, dtype=np.uint32) for v in (ct[0], ct[1], key[0], key[1])]
        rc = lib.or_mult_key_short_base(self._h, num_main, num_ter, *[_p(a) for a in ins],
                                        _p(ob), _p(oa))
        if rc != 0:
            raise ValueError("or_mult_key_short_base failed")
        return ob, oa

    def mod_up_to_max_ct(self, num_main, num_ter, ct, dts=None, std=None):
        """BootContext::ModUpToMax (src/extension/BootContext.cpp:186-268) of
        a ciphertext over the short base (1 or 2 primes) to the top level's
        q limbs; with dts / std keys (sparse-secret encapsulation) between
        the dense -> sparse switch at level -1 and the sparse -> dense switch
        at the top level. dts: 2 num_q limbs each; std: the top level's q
        limbs then alpha auxiliary limbs each."""
        sse = dts is not None
        tq = self.num_q(self.max_level)
        ob = np.empty((tq, self.n), dtype=np.uint32)
        oa = np.empty((tq, self.n), dtype=np.uint32)
        ins = [np.ascontiguousarray(v, dtype=np.uint32) for v in ct]
        keys = [np.ascontiguousarray(v, dtype=np.uint32) if sse else None
                for v in ((dts or (None, None)) + (std or (None, None)))]
        rc = lib.or_mod_up_to_max(self._h, num_main, num_ter, int(sse), _p(ins[0]), _p(ins[1]),
                                  *[_p(k) for k in keys], _p(ob), _p(oa))
        if rc != 0:
            raise ValueError("or_mod_up_to_max failed")
        return ob, oa

    def hconj(self, level, ct, ck):
        nq = self.num_q(level)
        ob = np.empty((nq, self.n), dtype=np.uint32)
        oa = np.empty((nq, self.n), dtype=np.uint32)
        ins = [np.ascontiguousarray(v, dtype=np.uint32) for v in (ct[0], ct[1], ck[0], ck[1])]
        rc = lib.or_hconj(self._h, level, *[_p(a) for a in ins], _p(ob), _p(oa))
        if rc != 0:
            raise ValueError("or_hconj failed")
        return ob, oa

    def hoisted_baby_step(self, level, ct, rk, rot):
        tot = self.num_q(level) + self.alpha
        ob = np.empty((tot, self.n), dtype=np.uint32)
        oa = np.empty((tot, self.n), dtype=np.uint32)
        ins = [np.ascontiguousarray(v, dtype=np.uint32) for v in (ct[0], ct[1], rk[0], rk[1])]
        rc = lib.or_hoisted_baby_step(self._h, level, *[_p(a) for a in ins], int(rot),
                                      _p(ob), _p(oa))
        if rc != 0:
            raise ValueError("or_hoisted_baby_step failed")
        return ob, oa

    def hrot(self, level, ct, rk, rot):
        nq = self.num_q(level)
        ob = np.empty((nq, self.n), dtype=np.uint32)
        oa = np.empty((nq, self.n), dtype=np.uint32)
        ins = [np.ascontiguousarray(v, dtype=np.uint32) for v in (ct[0], ct[1], rk[0], rk[1])]
        rc = lib.or_hrot(self._h, level, *[_p(a) for a in ins], int(rot), _p(ob), _p(oa))
        if rc != 0:
            raise ValueError("or_hrot failed")
        return ob, oa
