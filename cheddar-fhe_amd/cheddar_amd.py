"""Python binding of libcheddar_amd.so (the C ABI in include/cheddar_amd.h).

PyTorch provides device memory and streams here; every computation runs in
the HIP kernels of libcheddar_amd.so. Importing this module on a machine
without the built library raises: there is no CPU fallback.

Polynomials are torch int32 tensors on the GPU holding the u32 residues bit
for bit, shaped [limbs, N] (limb-major, the reference's layout).
"""
import ctypes
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcheddar_amd.so")
PARAM_DIR = os.path.join(HERE, "params")

CK_MODDOWN, CK_RESCALE, CK_MODDOWN_RESCALE = 0, 1, 2
CK_OPT_BCONV_MFMA = 1
CK_OPT_PLAN_AUTOTUNE = 2
CK_OPT_FUSED_KEY_ROWS = 3
CK_OPT_PLAN_NTT_FORM = 4
CK_OPT_PLAN_BCONV_FORM = 5
CK_OPT_NTT8_BLOCKS = 6
CK_OPT_NTT8_CHUNK = 7
# forms of transform / BaseConv plans (CK_OPT_PLAN_NTT_FORM, CK_OPT_PLAN_BCONV_FORM)
NTT_FORMS = {"valu": 0, "int8": 1, "auto": -1}
BCONV_FORMS = {"fp64": 0, "int8": 1, "valu": 2, "auto": -1}

_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_u32p = ctypes.POINTER(ctypes.c_uint32)

_SIGS = {
    "ck_last_error": (ctypes.c_char_p, []),
    "ck_version": (ctypes.c_int, []),
    "ck_context_create": (ctypes.c_int, [ctypes.c_int, _u32p, ctypes.c_int, _u32p, ctypes.c_int,
                                          _u32p, ctypes.c_int, _ip, ctypes.c_int,
                                          ctypes.POINTER(_vp)]),
    "ck_context_destroy": (ctypes.c_int, [_vp]),
    "ck_context_info": (ctypes.c_int, [_vp, _ip]),
    "ck_level_to_np": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip]),
    "ck_np_to_level": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "ck_modup_beta": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ck_context_set_short_base": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "ck_galois_factor": (ctypes.c_int, [_vp, ctypes.c_int, _u32p]),
    "ck_prime": (ctypes.c_int, [_vp, ctypes.c_int, _u32p]),
    "ck_ntt": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, _vp]),
    "ck_intt": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, _vp]),
    "ck_intt_mult_const": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, _u32p, ctypes.c_int, _vp]),
    "ck_modup": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp]),
    "ck_moddown": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]),
    "ck_add": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]),
    "ck_sub": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]),
    "ck_mult": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]),
    "ck_neg": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]),
    "ck_mult_const": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp, _vp]),
    "ck_add_const": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp, ctypes.c_int,
                                    _vp]),
    "ck_mult_imaginary_unit": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, ctypes.POINTER(_vp),
                                              ctypes.POINTER(_vp), _vp]),
    "ck_modup_to_max": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int,
                                       _vp]),
    "ck_crt_lift": (ctypes.c_int, [_vp, _ip, ctypes.c_int, _ip, ctypes.c_int, _vp, _vp, _vp]),
    "ck_paccum": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                 ctypes.POINTER(_vp), ctypes.c_int, _vp]),
    "ck_caccum": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                 ctypes.POINTER(_vp), ctypes.c_int, _vp]),
    "ck_tensor": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int] + [_vp] * 7 + [_vp]),
    "ck_permute": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int, _vp]),
    "ck_copy": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t, _vp]),
    "ck_mult_key_no_moddown": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                                              ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]),
    "ck_mult_key": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp,
                                   ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]),
    "ck_relinearize": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                                      ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int, _vp]),
    "ck_hmult": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int, _vp]),
    "ck_hmult_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                        ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int,
                                        _vp]),
    "ck_hrot": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp,
                               ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int, _vp]),
    "ck_hconj": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp,
                                ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]),
    "ck_rescale": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp),
                                  ctypes.POINTER(_vp), _vp]),
    "ck_hrot_hoisted": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _ip,
                                       ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp, _vp,
                                       ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int,
                                       _vp]),
    "ck_plan_hmult": (ctypes.c_int, [_vp, ctypes.c_int, _ip, ctypes.c_int] +
                      [ctypes.POINTER(_vp)] * 8 + [ctypes.POINTER(_vp)]),
    "ck_plan_hrot_hoisted": (ctypes.c_int, [_vp, ctypes.c_int, _ip, ctypes.POINTER(_vp),
                                            ctypes.POINTER(_vp), ctypes.c_int, _ip,
                                            ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                            ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                            ctypes.c_int, ctypes.POINTER(_vp)]),
    "ck_plan_run": (ctypes.c_int, [_vp, _vp]),
    "ck_plan_info": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_longlong)]),
    "ck_plan_destroy": (ctypes.c_int, [_vp]),
    "ck_plan_tuning": (ctypes.c_int, [_vp, ctypes.c_int, _ip, ctypes.POINTER(ctypes.c_double)]),
    "ck_context_set_option": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "ck_gs_accum": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                    ctypes.POINTER(_vp), _vp]),
    "ck_permute_accum": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.POINTER(_vp), ctypes.c_int,
                                         ctypes.POINTER(_vp), _ip, ctypes.c_int, _vp]),
    "ck_plan_bconv": (ctypes.c_int, [_vp, _ip, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                      ctypes.POINTER(_vp)]),
    "ck_plan_ntt": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp),
                                    ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "ck_profile_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ck_profile_report": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_char_p,
                                         ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_longlong)]),
}

EXPORTED = tuple(_SIGS)


def load_library(path=LIB_PATH):
    """Load libcheddar_amd.so and declare every entry point's signature."""
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run python cheddar-fhe_amd/build.py")
    # One HIP runtime per process: torch ships its own libamdhip64 (SONAME
    # libamdhip64.so.7, loaded by the name libamdhip64.so). Loading torch
    # first lets this library's libamdhip64.so.7 dependency bind to that
    # copy; the other order maps a second runtime and torch then finds no
    # device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load_library()


class CheddarError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise CheddarError(f"status {rc}: {lib.ck_last_error().decode()}")


def _ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def _ptrs(ts):
    arr = (ctypes.c_void_p * max(1, len(ts)))()
    for i, t in enumerate(ts):
        arr[i] = None if t is None else t.data_ptr()
    return arr


def _stream(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def load_param_json(path_or_name):
    p = path_or_name
    if not os.path.exists(p):
        p = os.path.join(PARAM_DIR, path_or_name)
    with open(p) as f:
        return json.load(f)


class Context:
    """Owns a ck_context: the reference's Parameter + Context pair
    (include/core/Parameter.h:42-47, include/core/Context.h:70)."""

    def __init__(self, log_degree, main_primes, aux_primes, ter_primes=(), level_config=None):
        main = (ctypes.c_uint32 * max(1, len(main_primes)))(*main_primes)
        ter = (ctypes.c_uint32 * max(1, len(ter_primes)))(*ter_primes)
        aux = (ctypes.c_uint32 * max(1, len(aux_primes)))(*aux_primes)
        if level_config is None:
            level_config = [[i + 1, 0] for i in range(len(main_primes))]
        flat = [int(v) for pair in level_config for v in pair]
        cfg = (ctypes.c_int * len(flat))(*flat)
        h = ctypes.c_void_p()
        _check(lib.ck_context_create(int(log_degree), main, len(main_primes),
                                     ter if len(ter_primes) else None, len(ter_primes),
                                     aux, len(aux_primes), cfg, len(level_config),
                                     ctypes.byref(h)))
        self._h = h
        info = (ctypes.c_int * 10)()
        _check(lib.ck_context_info(self._h, info))
        (self.log_n, self.n, self.max_main, self.max_ter, self.alpha, self.L, self.dnum,
         self.max_level, self.num_primes, self.workspace_mib) = list(info)
        self.level_config = [tuple(x) for x in level_config]

    @classmethod
    def from_json(cls, d):
        if isinstance(d, str):
            d = load_param_json(d)
        return cls(d["log_degree"], d["main_primes"], d["auxiliary_primes"],
                   d.get("terminal_primes", []), d["level_config"])

    def close(self):
        if getattr(self, "_h", None):
            lib.ck_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters -----------------------------------------------------
    def level_np(self, level):
        m, t = ctypes.c_int(), ctypes.c_int()
        _check(lib.ck_level_to_np(self._h, level, ctypes.byref(m), ctypes.byref(t)))
        return m.value, t.value

    def num_q(self, level):
        m, t = self.level_np(level)
        return m + t

    def np_to_level(self, num_main, num_ter):
        return lib.ck_np_to_level(self._h, num_main, num_ter)

    def modup_beta(self, level):
        return lib.ck_modup_beta(self._h, level)

    def set_short_base(self, num_main, num_ter):
        """level -1 of sparse-secret encapsulation: the short base's key
        switching with as many auxiliary primes (key blocks: the short base's
        limbs, then its auxiliary limbs)"""
        _check(lib.ck_context_set_short_base(self._h, num_main, num_ter))

    def galois_factor(self, rot):
        g = ctypes.c_uint32()
        _check(lib.ck_galois_factor(self._h, rot, ctypes.byref(g)))
        return g.value

    def prime(self, pidx):
        g = ctypes.c_uint32()
        _check(lib.ck_prime(self._h, pidx, ctypes.byref(g)))
        return g.value

    # ---- NTTHandler -----------------------------------------------------
    def ntt(self, dst, src, num_main, num_ter, num_aux=0, src_gap=0, mont_conv=False, stream=None):
        _check(lib.ck_ntt(self._h, _ptr(dst), _ptr(src), num_main, num_ter, num_aux, src_gap,
                          int(mont_conv), _stream(stream)))

    def intt(self, dst, src, num_main, num_ter, num_aux=0, src_gap=0, mont_conv=True, stream=None):
        _check(lib.ck_intt(self._h, _ptr(dst), _ptr(src), num_main, num_ter, num_aux, src_gap,
                           int(mont_conv), _stream(stream)))

    def intt_mult_const(self, dst, src, num_main, num_ter, num_aux, consts, normalize,
                        src_gap=0, stream=None):
        c = (ctypes.c_uint32 * len(consts))(*[int(v) for v in consts])
        _check(lib.ck_intt_mult_const(self._h, _ptr(dst), _ptr(src), num_main, num_ter, num_aux,
                                      src_gap, c, int(normalize), _stream(stream)))

    # ---- ModSwitchHandler -------------------------------------------------
    def modup(self, dst, src, level, stream=None):
        _check(lib.ck_modup(self._h, level, _ptr(dst), _ptr(src), _stream(stream)))

    def moddown(self, dsts, srcs, level, kind, stream=None):
        _check(lib.ck_moddown(self._h, level, kind, len(dsts), _ptrs(dsts), _ptrs(srcs),
                              _stream(stream)))

    # ---- ElementWiseHandler ----------------------------------------------
    def add(self, dsts, xs, ys, np3, stream=None):
        _check(lib.ck_add(self._h, *np3, len(dsts), _ptrs(dsts), _ptrs(xs), _ptrs(ys),
                          _stream(stream)))

    def sub(self, dsts, xs, ys, np3, stream=None):
        _check(lib.ck_sub(self._h, *np3, len(dsts), _ptrs(dsts), _ptrs(xs), _ptrs(ys),
                          _stream(stream)))

    def mult(self, dsts, xs, ys, np3, stream=None):
        _check(lib.ck_mult(self._h, *np3, len(dsts), _ptrs(dsts), _ptrs(xs), _ptrs(ys),
                           _stream(stream)))

    def neg(self, dsts, xs, np3, stream=None):
        _check(lib.ck_neg(self._h, *np3, len(dsts), _ptrs(dsts), _ptrs(xs), _stream(stream)))

    def mult_const(self, dsts, xs, consts_dev, np3, stream=None):
        _check(lib.ck_mult_const(self._h, *np3, len(dsts), _ptrs(dsts), _ptrs(xs),
                                 _ptr(consts_dev), _stream(stream)))

    def add_const(self, dsts, xs, consts_dev, np3, mode=0, stream=None):
        """mode 0: x + c, 1: x - c, 2: c - x (c per limb, device array)"""
        _check(lib.ck_add_const(self._h, *np3, len(dsts), _ptrs(dsts), _ptrs(xs),
                                _ptr(consts_dev), mode, _stream(stream)))

    def mult_imaginary_unit(self, dsts, xs, np3, stream=None):
        _check(lib.ck_mult_imaginary_unit(self._h, *np3, len(dsts), _ptrs(dsts), _ptrs(xs),
                                          _stream(stream)))

    def modup_to_max(self, dst, src, base_np, num_aux=0, stream=None):
        _check(lib.ck_modup_to_max(self._h, base_np[0], base_np[1], _ptr(src), _ptr(dst),
                                   num_aux, _stream(stream)))

    def paccum(self, dsts, xs, ys, np3, accumulate=False, stream=None):
        """dsts[p] (+)= sum_t xs[t][p] * ys[t]; xs: list of per-term lists of polys"""
        flat = [v for term in xs for v in term]
        _check(lib.ck_paccum(self._h, *np3, len(dsts), len(ys), _ptrs(dsts), _ptrs(flat),
                             _ptrs(ys), int(accumulate), _stream(stream)))

    def copy(self, dst, src, nbytes=None, stream=None):
        """device copy of nbytes bytes (default: all of src) from src to dst;
        both tensors must hold nbytes contiguous bytes"""
        have_s = src.numel() * src.element_size()
        have_d = dst.numel() * dst.element_size()
        n = have_s if nbytes is None else int(nbytes)
        if n < 0 or n > have_s or n > have_d or not (src.is_contiguous() and dst.is_contiguous()):
            raise ValueError(f"copy of {n} bytes between tensors of {have_s} and {have_d} "
                             "bytes (both contiguous)")
        _check(lib.ck_copy(self._h, _ptr(dst), _ptr(src), n, _stream(stream)))

    def gs_accum(self, dsts, xs, pts, np3, stream=None):
        """dsts[g] = [sum_t xs[t][j] * pts[g][t] for j in (b, a)] for every
        giant step g; xs: per-baby-step [b, a]; pts[g][t] may be None"""
        flat_d = [v for pair in dsts for v in pair]
        flat_x = [v for pair in xs for v in pair]
        flat_p = [v for row in pts for v in row]
        _check(lib.ck_gs_accum(self._h, *np3, len(xs), len(dsts), _ptrs(flat_d),
                               _ptrs(flat_x), _ptrs(flat_p), _stream(stream)))

    def permute_accum(self, dsts, srcs, rots, np3, accumulate=False, stream=None):
        """dsts[j] (+)= sum_s sigma_{rots[s]}(srcs[s][j]); srcs[s][j] may be None"""
        flat = [v for row in srcs for v in row]
        r = (ctypes.c_int * max(1, len(rots)))(*rots)
        _check(lib.ck_permute_accum(self._h, *np3, len(dsts), _ptrs(dsts), len(srcs),
                                    _ptrs(flat), r, int(accumulate), _stream(stream)))

    def caccum(self, dsts, xs, consts, np3, accumulate=False, stream=None):
        flat = [v for term in xs for v in term]
        _check(lib.ck_caccum(self._h, *np3, len(dsts), len(consts), _ptrs(dsts), _ptrs(flat),
                             _ptrs(consts), int(accumulate), _stream(stream)))

    def crt_lift(self, dst, src, src_primes, dst_primes, stream=None):
        sp = (ctypes.c_int * len(src_primes))(*src_primes)
        dp = (ctypes.c_int * len(dst_primes))(*dst_primes)
        _check(lib.ck_crt_lift(self._h, sp, len(src_primes), dp, len(dst_primes), _ptr(src),
                               _ptr(dst), _stream(stream)))

    def tensor(self, d0, d1, d2, b1, a1, b2, a2, num_main, num_ter, stream=None):
        _check(lib.ck_tensor(self._h, num_main, num_ter, _ptr(d0), _ptr(d1), _ptr(d2),
                             _ptr(b1), _ptr(a1), _ptr(b2), _ptr(a2), _stream(stream)))

    def permute(self, dsts, xs, rot, np3, stream=None):
        _check(lib.ck_permute(self._h, *np3, len(dsts), _ptrs(dsts), _ptrs(xs), rot,
                              _stream(stream)))

    # ---- Context evaluator -------------------------------------------------
    def mult_key_no_moddown(self, acc_b, acc_a, x, key, level, add_b=None, add_a=None,
                            stream=None):
        _check(lib.ck_mult_key_no_moddown(self._h, level, _ptr(acc_b), _ptr(acc_a), _ptr(x),
                                          _ptr(add_b), _ptr(add_a), _ptrs(key[0]), _ptrs(key[1]),
                                          _stream(stream)))

    def mult_key(self, out, ct, key, level, stream=None):
        _check(lib.ck_mult_key(self._h, level, _ptr(out[0]), _ptr(out[1]), _ptr(ct[0]),
                               _ptr(ct[1]), _ptrs(key[0]), _ptrs(key[1]), _stream(stream)))

    def relinearize(self, out, d, key, level, rescale=True, stream=None):
        _check(lib.ck_relinearize(self._h, level, _ptr(out[0]), _ptr(out[1]), _ptr(d[0]),
                                  _ptr(d[1]), _ptr(d[2]), _ptrs(key[0]), _ptrs(key[1]),
                                  int(rescale), _stream(stream)))

    def hmult(self, out, ct1, ct2, key, level, rescale=True, stream=None):
        _check(lib.ck_hmult(self._h, level, _ptr(out[0]), _ptr(out[1]), _ptr(ct1[0]),
                            _ptr(ct1[1]), _ptr(ct2[0]), _ptr(ct2[1]), _ptrs(key[0]),
                            _ptrs(key[1]), int(rescale), _stream(stream)))

    def hmult_noalloc(self, out, ct1, ct2, key, level, rescale=True, stream=None):
        """hmult with its temporaries in the library's static device array
        (ck_hmult_noalloc: allocates nothing; calls on one device must not
        overlap)"""
        _check(lib.ck_hmult_noalloc(self._h, level, _ptr(out[0]), _ptr(out[1]),
                                    _ptr(ct1[0]), _ptr(ct1[1]), _ptr(ct2[0]), _ptr(ct2[1]),
                                    _ptrs(key[0]), _ptrs(key[1]), int(rescale),
                                    _stream(stream)))

    def hrot(self, out, ct, key, rot, level, stream=None):
        _check(lib.ck_hrot(self._h, level, _ptr(out[0]), _ptr(out[1]), _ptr(ct[0]), _ptr(ct[1]),
                           _ptrs(key[0]), _ptrs(key[1]), int(rot), _stream(stream)))

    def hconj(self, out, ct, key, level, stream=None):
        _check(lib.ck_hconj(self._h, level, _ptr(out[0]), _ptr(out[1]), _ptr(ct[0]), _ptr(ct[1]),
                            _ptrs(key[0]), _ptrs(key[1]), _stream(stream)))

    def rescale(self, dsts, srcs, level, stream=None):
        _check(lib.ck_rescale(self._h, level, len(dsts), _ptrs(dsts), _ptrs(srcs),
                              _stream(stream)))

    def hrot_hoisted(self, outs, ct, keys, rots, level, moddown=True, stream=None):
        """Hoisted rotations of ct = (b, a) by each of rots (in [1, N/2)):
        outs[i] = (b_i, a_i) receives HRot(ct, rots[i]) (moddown) or the
        Q*P-basis baby-step output; keys[i] = (key_b blocks, key_a blocks)."""
        n = len(rots)
        kb = [t for k in keys for t in k[0]]
        ka = [t for k in keys for t in k[1]]
        _check(lib.ck_hrot_hoisted(self._h, level, n, (ctypes.c_int * n)(*rots),
                                   _ptrs([o[0] for o in outs]), _ptrs([o[1] for o in outs]),
                                   _ptr(ct[0]), _ptr(ct[1]), _ptrs(kb), _ptrs(ka),
                                   int(moddown), _stream(stream)))

    def set_plan_autotune(self, on=True):
        """Batched plans built in both BaseConv forms, timed, faster kept."""
        _check(lib.ck_context_set_option(self._h, CK_OPT_PLAN_AUTOTUNE, int(bool(on))))

    def set_bconv_mfma(self, on=True):
        """BaseConv on the FP64 matrix cores (True), on the int8 matrix cores
        ("int8"), fused into the forward NTT column pass (False), or None for
        the default (int8 matrix cores in batched plans, fused in single
        operations); outputs are identical."""
        v = -1 if on is None else 2 if on == "int8" else int(bool(on))
        _check(lib.ck_context_set_option(self._h, CK_OPT_BCONV_MFMA, v))

    def set_plan_forms(self, ntt=None, bconv=None, ntt8_blocks=None, ntt8_chunk=None):
        """Forms of ck_plan_ntt forward transforms ("valu", "int8", "auto")
        and of ck_plan_bconv ("fp64", "int8", "valu", "auto"), and the int8
        transform's blocks per workgroup / chunk; outputs are identical."""
        if ntt is not None:
            _check(lib.ck_context_set_option(self._h, CK_OPT_PLAN_NTT_FORM, NTT_FORMS[ntt]))
        if bconv is not None:
            _check(lib.ck_context_set_option(self._h, CK_OPT_PLAN_BCONV_FORM,
                                             BCONV_FORMS[bconv]))
        if ntt8_blocks is not None:
            _check(lib.ck_context_set_option(self._h, CK_OPT_NTT8_BLOCKS, int(ntt8_blocks)))
        if ntt8_chunk is not None:
            _check(lib.ck_context_set_option(self._h, CK_OPT_NTT8_CHUNK, int(ntt8_chunk)))

    def set_fused_key_rows(self, on=True):
        """Key switches run the ModUp's forward row stages and the key inner
        product as one pass (True, default) or as a row pass plus a separate
        accumulation (False); outputs are identical."""
        _check(lib.ck_context_set_option(self._h, CK_OPT_FUSED_KEY_ROWS, int(bool(on))))

    # ---- per-kernel timing --------------------------------------------------
    def profile(self, on=True):
        _check(lib.ck_profile_enable(self._h, int(on)))

    def profile_report(self, max_entries=64):
        """{kernel name: (total ms, launches)} of the launches recorded since
        profile(True)"""
        names = ctypes.create_string_buffer(64 * max_entries)
        ms = (ctypes.c_double * max_entries)()
        calls = (ctypes.c_longlong * max_entries)()
        n = lib.ck_profile_report(self._h, max_entries, names, ms, calls)
        if n < 0:
            raise CheddarError(lib.ck_last_error().decode())
        raw = names.raw
        out = {}
        for i in range(n):
            name = raw[64 * i:64 * (i + 1)].split(b"\0", 1)[0].decode()
            out[name] = (ms[i], calls[i])
        return out


class Plan:
    """A batched plan (ck_plan_*): built once over fixed device buffers, run
    any number of times; keeps references to the tensors it was built on."""

    def __init__(self, ctx, handle, keep):
        self._ctx = ctx
        self._h = handle
        self._keep = keep

    @classmethod
    def hmult(cls, ctx, levels, outs, ct1s, ct2s, key, rescale=True):
        n = len(levels)
        h = ctypes.c_void_p()
        _check(lib.ck_plan_hmult(ctx._h, n, (ctypes.c_int * n)(*levels), int(rescale),
                                 _ptrs([o[0] for o in outs]), _ptrs([o[1] for o in outs]),
                                 _ptrs([c[0] for c in ct1s]), _ptrs([c[1] for c in ct1s]),
                                 _ptrs([c[0] for c in ct2s]), _ptrs([c[1] for c in ct2s]),
                                 _ptrs(key[0]), _ptrs(key[1]), ctypes.byref(h)))
        return cls(ctx, h, (outs, ct1s, ct2s, key))

    @classmethod
    def hrot_hoisted(cls, ctx, levels, cts, keys, rots, outs, moddown=True):
        """outs[i][r] = (b, a) output of rotation rots[r] of cts[i]"""
        n, nr = len(levels), len(rots)
        kb = [t for k in keys for t in k[0]]
        ka = [t for k in keys for t in k[1]]
        flat = [o for row in outs for o in row]
        h = ctypes.c_void_p()
        _check(lib.ck_plan_hrot_hoisted(ctx._h, n, (ctypes.c_int * n)(*levels),
                                        _ptrs([c[0] for c in cts]), _ptrs([c[1] for c in cts]),
                                        nr, (ctypes.c_int * nr)(*rots), _ptrs(kb), _ptrs(ka),
                                        _ptrs([o[0] for o in flat]), _ptrs([o[1] for o in flat]),
                                        int(moddown), ctypes.byref(h)))
        return cls(ctx, h, (cts, keys, outs))

    @classmethod
    def bconv(cls, ctx, src_primes, dst_primes, srcs, dsts):
        """dsts[k] = BaseConv of the centred int32 limbs srcs[k] from
        src_primes to dst_primes (ck_plan_bconv)"""
        n = len(srcs)
        h = ctypes.c_void_p()
        _check(lib.ck_plan_bconv(ctx._h, (ctypes.c_int * len(src_primes))(*src_primes),
                                 len(src_primes), (ctypes.c_int * len(dst_primes))(*dst_primes),
                                 len(dst_primes), n, _ptrs(srcs), _ptrs(dsts), ctypes.byref(h)))
        return cls(ctx, h, (srcs, dsts))

    @classmethod
    def ntt(cls, ctx, srcs, dsts, np3, inverse=False, mont_conv=False):
        """dsts[k] = NTT (or INTT) of srcs[k] over the limbs of np3 (ck_plan_ntt)"""
        h = ctypes.c_void_p()
        _check(lib.ck_plan_ntt(ctx._h, len(srcs), *np3, int(inverse), int(mont_conv),
                               _ptrs(srcs), _ptrs(dsts), ctypes.byref(h)))
        return cls(ctx, h, (srcs, dsts))

    def run(self, stream=None):
        _check(lib.ck_plan_run(self._h, _stream(stream)))

    def info(self):
        out = (ctypes.c_longlong * 8)()
        _check(lib.ck_plan_info(self._h, out))
        # form: a transform / BaseConv plan's form, or an HMult plan's key
        # pass (1: ModUp row stages fused with the key product, 0: separate)
        return {"items": out[0], "steps": out[1], "launches": out[2],
                "workspace_bytes": out[3], "bconv_mfma": bool(out[4]),
                "bconv_form": ("fused", "fp64", "int8")[out[4]] if out[4] < 3 else None,
                "tuned_us": (out[5] / 1e3, out[6] / 1e3) if out[5] != -1000 else None,
                "form": out[7]}

    def tuning(self):
        """[(form, int8 blocks, int8 chunk, us)] of every autotuned candidate of
        a transform / BaseConv plan, [(BaseConv form, key pass, item chunks,
        us)] of an HMult / hoisted plan"""
        k = lib.ck_plan_tuning(self._h, 0, None, None)
        if k < 0:
            raise CheddarError(lib.ck_last_error().decode())
        if k == 0:
            return []
        forms = (ctypes.c_int * (3 * k))()
        us = (ctypes.c_double * k)()
        lib.ck_plan_tuning(self._h, k, forms, us)
        return [(forms[3 * i], forms[3 * i + 1], forms[3 * i + 2], us[i]) for i in range(k)]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and lib is not None:
            lib.ck_plan_destroy(h)
            self._h = None


# ---- host <-> device helpers ------------------------------------------------
def to_device(arr_u32, device="cuda"):
    """numpy uint32 [limbs, N] -> torch int32 device tensor (same bits)."""
    import numpy as np
    import torch
    a = np.ascontiguousarray(arr_u32, dtype=np.uint32).view(np.int32)
    return torch.from_numpy(a).to(device)


def to_host(t):
    """torch int32 device tensor -> numpy uint32 (same bits)."""
    import numpy as np
    return t.detach().cpu().numpy().view(np.uint32)
