"""The C ABI library loads without a GPU and exports every declared symbol."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "onionpir_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(opir_[a-z0-9_]+)\s*\(", src)))


def test_header_matches_binding_list():
    from onionpirv2_amd import gpu
    assert declared_functions() == sorted(gpu.EXPORTED)


def test_library_exports_every_symbol():
    from onionpirv2_amd import gpu
    lib = gpu.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_ctx_create_rejects_bad_scheme_without_gpu():
    """Argument validation happens before any device call."""
    from onionpirv2_amd import gpu
    lib = gpu.load_library()
    s = gpu.OpirScheme(poly_degree=4096, composite=1)
    h = ctypes.c_void_p()
    assert lib.opir_ctx_create(ctypes.byref(s), 0, ctypes.byref(h)) == gpu.OPIR_EINVAL
    assert b"composite" in lib.opir_last_error(None)
    assert lib.opir_answer(None, 0, None, None) == gpu.OPIR_EINVAL


def test_32bit_limb_primes_above_2_29_rejected_without_gpu(O):
    """The 32-bit kernels take limb primes below 2^29 (the forward transform
    lets values grow to 8q < 2^32): two primes 1 mod 2N in [2^29, 2^30) are
    rejected before any device call, composite or as two RNS limbs."""
    from onionpirv2_amd import gpu
    from conftest import scheme_from_oracle
    from test_oracle_pins_k2 import _is_prime
    lib = gpu.load_library()
    p = O.PirParams("CONFIG_N2048_K1_COMP")
    m = 2 * p.N
    big = [q for q in range((1 << 29) // m * m + 1, 1 << 30, m) if q > (1 << 29) and _is_prime(q)][:2]
    assert len(big) == 2 and all((1 << 29) < q < (1 << 30) and q % m == 1 for q in big)
    for composite in (1, 0):
        s = gpu.OpirScheme(**{**scheme_from_oracle(p), "composite": composite, "q1": big[0], "q2": big[1]})
        h = ctypes.c_void_p()
        assert lib.opir_ctx_create(ctypes.byref(s), 0, ctypes.byref(h)) == gpu.OPIR_EINVAL
        assert b"2^29" in lib.opir_last_error(None)


def test_k2_scheme_validation_without_gpu(O):
    """The two-prime scheme of CONFIG_N4096_K2_MP passes host-side validation
    (it then needs a device); malformed variants are rejected first."""
    from onionpirv2_amd import gpu
    lib = gpu.load_library()
    p = O.PirParams("CONFIG_N4096_K2_MP")
    q0, q1 = p.rns_mods
    good = dict(poly_degree=p.N, composite=0, q1=q0, q2=q1, ntt_root=O.ntt_root(p.N, q0),
                ntt_root2=O.ntt_root(p.N, q1), plain_mod=p.plain_mod, small_q=p.small_q, l_ep=p.l_ep,
                base_log2=p.base_log2, l_key=p.l_key, base_log2_key=p.base_log2_key, l_ks=p.l_ks,
                bv_base_log2=p.bv_base_log2)

    def create(**over):
        s = gpu.OpirScheme(**{**good, **over})
        h = ctypes.c_void_p()
        rc = lib.opir_ctx_create(ctypes.byref(s), 0, ctypes.byref(h))
        if rc == gpu.OPIR_OK:
            lib.opir_ctx_destroy(h)
        return rc

    assert create() in (gpu.OPIR_OK, gpu.OPIR_EHIP)       # valid: fails only for want of a device
    assert create(ntt_root2=2) == gpu.OPIR_EINVAL          # not a primitive 2N-th root
    assert create(bv_base_log2=20) == gpu.OPIR_EINVAL      # last signed digit shift beyond 126 bits
    assert create(small_q=q0 + 1) == gpu.OPIR_EINVAL
    assert create(poly_degree=2048) == gpu.OPIR_EINVAL
    # a 128-bit multiply-accumulate sum holds at most 64 products of
    # residues (< 4q times < q, q < 2^60): 2 * l_ep and 2 * l_key <= 64
    assert create(l_ep=33, base_log2=3) == gpu.OPIR_EINVAL
    assert create(l_key=33, base_log2_key=3) == gpu.OPIR_EINVAL
    assert create(l_ep=32, base_log2=3) in (gpu.OPIR_OK, gpu.OPIR_EHIP)


def test_single_prime_scheme_validation_without_gpu(O):
    """CONFIG_N2048_K1 (one ~60-bit prime, q2 = 0) passes host-side
    validation; malformed variants are rejected before any device call."""
    from onionpirv2_amd import gpu
    lib = gpu.load_library()
    p = O.PirParams("CONFIG_N2048_K1")
    (q,) = p.rns_mods
    good = dict(poly_degree=p.N, composite=0, q1=q, q2=0, ntt_root=O.ntt_root(p.N, q), plain_mod=p.plain_mod,
                small_q=p.small_q, l_ep=p.l_ep, base_log2=p.base_log2, l_key=p.l_key, base_log2_key=p.base_log2_key,
                l_ks=p.l_ks, bv_base_log2=p.bv_base_log2)

    def create(**over):
        s = gpu.OpirScheme(**{**good, **over})
        h = ctypes.c_void_p()
        rc = lib.opir_ctx_create(ctypes.byref(s), 0, ctypes.byref(h))
        if rc == gpu.OPIR_OK:
            lib.opir_ctx_destroy(h)
        return rc

    assert create() in (gpu.OPIR_OK, gpu.OPIR_EHIP)     # valid: fails only for want of a device
    assert create(ntt_root=2) == gpu.OPIR_EINVAL         # not a primitive 2N-th root
    assert create(poly_degree=4096) == gpu.OPIR_EINVAL   # one prime runs at N = 2048
    assert create(small_q=q) == gpu.OPIR_EINVAL
    # balanced digits over a centred int64: the last offset 2^(b(l-1)-1) must fit
    assert create(base_log2=16) == gpu.OPIR_EINVAL       # 16 * 4 = 64 > 62
    assert create(base_log2=15) in (gpu.OPIR_OK, gpu.OPIR_EHIP)
    assert create(bv_base_log2=9) == gpu.OPIR_EINVAL     # 9 * 7 = 63 > 62
