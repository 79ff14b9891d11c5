"""ctypes binding of the C ABI in include/onionpir_gpu.h.

This is the boundary the GPU parity tests and bench.py call. Loading fails
loudly when libonionpir_gpu.so has not been built: there is no CPU fallback.
"""
import ctypes
import weakref
import os

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libonionpir_gpu.so")

OPIR_OK = 0
OPIR_EINVAL = -1
OPIR_ERANGE = -2
OPIR_ERUNTIME = -3
OPIR_EHIP = -4
OPIR_ENOMEM = -5
SCAN_DEFAULT = -1
SCAN_AUTOTUNE = -2

# Every symbol include/onionpir_gpu.h declares.
EXPORTED = [
    "opir_ctx_create", "opir_ctx_destroy", "opir_last_error", "opir_ctx_set_stream", "opir_ctx_synchronize",
    "opir_ctx_clone", "opir_db_view",
    "opir_tree_create", "opir_tree_destroy", "opir_tree_shape", "opir_tree_load_level", "opir_tree_select",
    "opir_db_create", "opir_db_destroy", "opir_db_load_plain", "opir_db_load_ntt", "opir_db_load_plain_u64", "opir_db_gen_random",
    "opir_client_set_keys", "opir_client_set_keys_seeded", "opir_client_remove", "opir_client_get_keys",
    "opir_client_keygen_device", "opir_query_gen", "opir_query_gen_device",
    "opir_answer", "opir_answer_batch", "opir_answer_batch_device",
    "opir_db_workspace_bytes", "opir_db_reserve", "opir_db_trim", "opir_client_key_bytes", "opir_ctx_resident_key_bytes",
    "opir_ctx_set_profiling", "opir_ctx_set_cu_mask", "opir_last_stage_ms", "opir_ctx_set_scan_variant", "opir_ctx_set_graphs", "opir_scan_variant_count", "opir_db_tuned_scan_variant",
    "opir_ntt_fwd", "opir_ntt_inv", "opir_expand", "opir_complete", "opir_first_dim", "opir_other_dim",
    "opir_ext_prod", "opir_ext_prod_mux", "opir_mod_switch", "opir_bv_ks", "opir_ext_prod_mux_device_noalloc",
]


class InvalidArgument(ValueError):
    """OPIR_EINVAL: the reference throws std::invalid_argument."""


class OutOfRange(IndexError):
    """OPIR_ERANGE: the reference throws std::out_of_range."""


class OpirRuntimeError(RuntimeError):
    """Any other failure (std::runtime_error in the reference)."""


class OpirScheme(ctypes.Structure):
    _fields_ = [
        ("poly_degree", ctypes.c_uint32), ("composite", ctypes.c_uint32),
        ("q1", ctypes.c_uint64), ("q2", ctypes.c_uint64), ("ntt_root", ctypes.c_uint64),
        ("plain_mod", ctypes.c_uint64), ("small_q", ctypes.c_uint64),
        ("l_ep", ctypes.c_uint32), ("base_log2", ctypes.c_uint32),
        ("l_key", ctypes.c_uint32), ("base_log2_key", ctypes.c_uint32),
        ("l_ks", ctypes.c_uint32), ("bv_base_log2", ctypes.c_uint32),
        ("ntt_root2", ctypes.c_uint64),
        ("noise_std_dev", ctypes.c_double),
    ]


_lib = None


def load_library():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise ImportError(f"{_LIB_PATH} is not built; run `python -m onionpirv2_amd.build gpu`")
    lib = ctypes.CDLL(_LIB_PATH)
    P = ctypes.c_void_p
    u64p = ctypes.POINTER(ctypes.c_uint64)
    sz = ctypes.c_size_t
    sig = {
        "opir_ctx_create": ([ctypes.POINTER(OpirScheme), ctypes.c_int, ctypes.POINTER(P)], ctypes.c_int),
        "opir_ctx_destroy": ([P], None),
        "opir_last_error": ([P], ctypes.c_char_p),
        "opir_ctx_set_stream": ([P, P], ctypes.c_int),
        "opir_ctx_synchronize": ([P], ctypes.c_int),
        "opir_ctx_clone": ([P, ctypes.POINTER(P)], ctypes.c_int),
        "opir_db_view": ([P, P, ctypes.POINTER(P)], ctypes.c_int),
        "opir_tree_create": ([P, sz, sz, ctypes.POINTER(P)], ctypes.c_int),
        "opir_tree_destroy": ([P], None),
        "opir_tree_shape": ([P] + [ctypes.POINTER(sz)] * 4, ctypes.c_int),
        "opir_tree_load_level": ([P, sz, P], ctypes.c_int),
        "opir_tree_select": ([P, ctypes.c_uint64, P, P], ctypes.c_int),
        "opir_db_create": ([P, sz, sz, sz, sz, ctypes.POINTER(P)], ctypes.c_int),
        "opir_db_destroy": ([P], None),
        "opir_db_load_plain": ([P, P, sz, sz], ctypes.c_int),
        "opir_db_load_ntt": ([P, P, P], ctypes.c_int),
        "opir_db_load_plain_u64": ([P, P, sz, sz], ctypes.c_int),
        "opir_db_gen_random": ([P, ctypes.c_uint64], ctypes.c_int),
        "opir_client_set_keys": ([P, ctypes.c_uint64, P, sz, P], ctypes.c_int),
        "opir_client_remove": ([P, ctypes.c_uint64], ctypes.c_int),
        "opir_client_set_keys_seeded": ([P, ctypes.c_uint64, P, P, sz, P], ctypes.c_int),
        "opir_client_get_keys": ([P, ctypes.c_uint64, P, sz, P], ctypes.c_int),
        "opir_client_keygen_device": ([P, sz, P, P, sz], ctypes.c_int),
        "opir_query_gen": ([P, sz, P, P, P, P], ctypes.c_int),
        "opir_query_gen_device": ([P, sz, P, P, P, P], ctypes.c_int),
        "opir_answer": ([P, ctypes.c_uint64, P, P], ctypes.c_int),
        "opir_answer_batch": ([P, sz, P, P, P], ctypes.c_int),
        "opir_answer_batch_device": ([P, sz, P, P, P], ctypes.c_int),
        "opir_db_workspace_bytes": ([P, sz, ctypes.POINTER(sz)], ctypes.c_int),
        "opir_db_reserve": ([P, sz], ctypes.c_int),
        "opir_db_trim": ([P, sz], ctypes.c_int),
        "opir_client_key_bytes": ([P, sz, ctypes.POINTER(sz)], ctypes.c_int),
        "opir_ctx_resident_key_bytes": ([P, ctypes.POINTER(sz)], ctypes.c_int),
        "opir_ctx_set_profiling": ([P, ctypes.c_int], ctypes.c_int),
        "opir_last_stage_ms": ([P, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
        "opir_ctx_set_scan_variant": ([P, ctypes.c_int], ctypes.c_int),
        "opir_ctx_set_cu_mask": ([P, P, sz], ctypes.c_int),
        "opir_ctx_set_graphs": ([P, ctypes.c_int], ctypes.c_int),
        "opir_scan_variant_count": ([P, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "opir_db_tuned_scan_variant": ([P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "opir_ntt_fwd": ([P, P, sz], ctypes.c_int),
        "opir_ntt_inv": ([P, P, sz], ctypes.c_int),
        "opir_expand": ([P, ctypes.c_uint64, P, P], ctypes.c_int),
        "opir_complete": ([P, ctypes.c_uint64, P, P], ctypes.c_int),
        "opir_first_dim": ([P, P, P], ctypes.c_int),
        "opir_other_dim": ([P, P, P, P], ctypes.c_int),
        "opir_ext_prod": ([P, P, sz, sz, P, P], ctypes.c_int),
        "opir_ext_prod_mux": ([P, P, sz, sz, P, P, P], ctypes.c_int),
        "opir_mod_switch": ([P, P, P], ctypes.c_int),
        "opir_bv_ks": ([P, P, ctypes.c_uint32, P, P], ctypes.c_int),
        "opir_ext_prod_mux_device_noalloc": ([P, P, sz, sz, P, P, P], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    del u64p
    _lib = lib
    return lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _u64(a, shape=None):
    """Contiguous uint64 copy/view of a; with `shape`, the element count must
    match it exactly (the C ABI reads exactly that many values)."""
    if isinstance(a, np.ndarray) and a.dtype.kind == "i" and a.size and int(a.min()) < 0:
        raise InvalidArgument("negative value where a residue is expected")
    a = np.ascontiguousarray(a, dtype=np.uint64)
    if shape is not None:
        shape = tuple(int(x) for x in shape)
        if a.size != int(np.prod(shape)):
            raise InvalidArgument(f"array of {a.size} values where {shape} is expected")
        a = a.reshape(shape)
    return a


class GpuContext:
    """One opir_ctx: scheme tables, client keys and a HIP stream on one device."""

    def __init__(self, scheme: dict, device: int = 0, _parent=None):
        self.lib = load_library()
        self.scheme = dict(scheme)
        self.N = scheme["poly_degree"]
        # RNS limbs per polynomial at the boundary: 1 for the composite scheme and
        # for one prime (q2 = 0), else 2
        self.K = 1 if scheme.get("composite", 1) or not scheme.get("q2", 0) else 2
        self.KN = self.K * self.N
        h = ctypes.c_void_p()
        if _parent is None:
            s = OpirScheme(**scheme)
            rc = self.lib.opir_ctx_create(ctypes.byref(s), device, ctypes.byref(h))
            if rc != OPIR_OK:
                self._raise(rc, self.lib.opir_last_error(None))
        else:
            _parent.check(self.lib.opir_ctx_clone(_parent.h, ctypes.byref(h)))
        self.h = h
        self._parent = _parent  # a clone keeps its parent alive
        self._dbs = weakref.WeakSet()  # databases to destroy before the context

    def clone(self):
        """A context with this one's scheme and client keys but its own
        streams (opir_ctx_clone)."""
        return GpuContext(self.scheme, _parent=self)

    def _raise(self, rc, msg):
        msg = msg.decode() if isinstance(msg, bytes) else str(msg)
        if rc == OPIR_EINVAL:
            raise InvalidArgument(msg)
        if rc == OPIR_ERANGE:
            raise OutOfRange(msg)
        raise OpirRuntimeError(f"opir error {rc}: {msg}")

    def check(self, rc):
        if rc != OPIR_OK:
            self._raise(rc, self.lib.opir_last_error(self.h))

    def close(self):
        if getattr(self, "h", None):
            for db in list(getattr(self, "_dbs", ())):
                db.close()
            self.lib.opir_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, hip_stream_ptr):
        self.check(self.lib.opir_ctx_set_stream(self.h, ctypes.c_void_p(hip_stream_ptr or 0)))

    def synchronize(self):
        self.check(self.lib.opir_ctx_synchronize(self.h))

    def set_profiling(self, on=True):
        self.check(self.lib.opir_ctx_set_profiling(self.h, 1 if on else 0))

    def set_scan_variant(self, variant=-1):
        """First-dimension scan tile shape: SCAN_DEFAULT (-1, table per batch
        size), SCAN_AUTOTUNE (-2, timed on first use per batch-size class), or
        a forced shape index."""
        self.check(self.lib.opir_ctx_set_scan_variant(self.h, int(variant)))

    def set_cu_mask(self, cus=None):
        """Run this context's work on the compute units listed in `cus`
        (indices); None restores unrestricted streams."""
        if cus is None:
            self.check(self.lib.opir_ctx_set_cu_mask(self.h, None, 0))
            return
        cus = list(cus)
        words = np.zeros(max(cus, default=0) // 32 + 1, dtype=np.uint32)
        for c in cus:
            words[c // 32] |= np.uint32(1 << (c % 32))
        self.check(self.lib.opir_ctx_set_cu_mask(self.h, words.ctypes.data_as(ctypes.c_void_p), len(words)))

    def set_graphs(self, on=True):
        """Capture each answer's launch sequence as a HIP graph on first use
        (per batch size, key-table slot, buffers and scan shape) and replay it
        afterwards."""
        self.check(self.lib.opir_ctx_set_graphs(self.h, 1 if on else 0))

    def scan_variant_count(self):
        out = ctypes.c_int(0)
        self.check(self.lib.opir_scan_variant_count(self.h, ctypes.byref(out)))
        return out.value

    def last_stage_ms(self):
        out = (ctypes.c_float * 6)()
        self.check(self.lib.opir_last_stage_ms(self.h, out))
        names = ["Expand", "Convert", "First dim", "Other dim", "Modulus switching", "scan kernel"]
        return dict(zip(names, [float(x) for x in out]))

    def set_client_keys(self, client_id, bv, gsw):
        sc = self.scheme
        bv = np.asarray(bv)
        h = bv.shape[0] if bv.ndim else 0
        bv = _u64(bv, (h, sc["l_ks"], 2, self.KN))
        gsw = _u64(gsw, (2 * sc["l_key"], 2 * self.KN))
        self.check(self.lib.opir_client_set_keys(self.h, client_id, _ptr(bv), bv.shape[0], _ptr(gsw)))

    def set_client_keys_seeded(self, client_id, seed, bv_b, gsw_c0):
        """Seed-compressed keys: bv_b [h, l_ks, K*N] (b halves), gsw_c0
        [2 l_key, K*N] (c0 halves), NTT form; the uniform halves are
        regenerated on the device from the 32-byte seed."""
        sc = self.scheme
        seed = bytes(seed)
        if len(seed) != 32:
            raise InvalidArgument("the key seed is 32 bytes")
        bv_b = np.asarray(bv_b)
        h = bv_b.shape[0] if bv_b.ndim else 0
        b = _u64(bv_b, (h, sc["l_ks"], self.KN))
        g = _u64(gsw_c0, (2 * sc["l_key"], self.KN))
        sb = (ctypes.c_uint8 * 32).from_buffer_copy(seed)
        self.check(self.lib.opir_client_set_keys_seeded(self.h, client_id, ctypes.cast(sb, ctypes.c_void_p),
                                                        _ptr(b), h, _ptr(g)))

    def get_client_keys(self, client_id, h):
        """The resident keys in set_client_keys's layout: (bv [h, l_ks, 2, K*N], gsw [2 l_key, 2*K*N])."""
        sc = self.scheme
        bv = np.zeros((h, sc["l_ks"], 2, self.KN), np.uint64)
        gsw = np.zeros((2 * sc["l_key"], 2 * self.KN), np.uint64)
        self.check(self.lib.opir_client_get_keys(self.h, client_id, _ptr(bv), h, _ptr(gsw)))
        return bv, gsw

    def keygen_device(self, client_ids, keys, h):
        """Generate and install the keys of seeded clients on the device
        (keys: one 32-byte seed per client)."""
        ids = np.ascontiguousarray(client_ids, dtype=np.uint64)
        kb = np.frombuffer(b"".join(bytes(k) for k in keys), dtype=np.uint8).copy()
        if kb.size != 32 * len(ids):
            raise InvalidArgument("one 32-byte key per client")
        self.check(self.lib.opir_client_keygen_device(self.h, len(ids), _ptr(ids), _ptr(kb), h))

    def key_bytes(self, h):
        """Device bytes of one client's keys at expansion height h."""
        out = ctypes.c_size_t(0)
        self.check(self.lib.opir_client_key_bytes(self.h, h, ctypes.byref(out)))
        return out.value

    def resident_key_bytes(self):
        """Device bytes of all client keys resident on this context."""
        out = ctypes.c_size_t(0)
        self.check(self.lib.opir_ctx_resident_key_bytes(self.h, ctypes.byref(out)))
        return out.value

    def remove_client(self, client_id):
        self.check(self.lib.opir_client_remove(self.h, client_id))

    def ntt_fwd(self, polys):
        a = _u64(polys).copy()
        self.check(self.lib.opir_ntt_fwd(self.h, _ptr(a), a.size // self.KN))
        return a

    def ntt_inv(self, polys):
        a = _u64(polys).copy()
        self.check(self.lib.opir_ntt_inv(self.h, _ptr(a), a.size // self.KN))
        return a

    def ext_prod(self, gsw, l, base_log2, ct):
        g = _u64(gsw, (2 * l, 2 * self.KN))
        c = _u64(ct, (2, self.KN))
        out = np.zeros((2, self.KN), dtype=np.uint64)
        self.check(self.lib.opir_ext_prod(self.h, _ptr(g), l, base_log2, _ptr(c), _ptr(out)))
        return out

    def ext_prod_mux(self, gsw, l, base_log2, x, y):
        g, xa, ya = _u64(gsw, (2 * l, 2 * self.KN)), _u64(x, (2, self.KN)), _u64(y, (2, self.KN))
        out = np.zeros((2, self.KN), dtype=np.uint64)
        self.check(self.lib.opir_ext_prod_mux(self.h, _ptr(g), l, base_log2, _ptr(xa), _ptr(ya), _ptr(out)))
        return out

    def bv_ks(self, ct, galois_k, key_rows):
        """bv_apply_galois_inplace: key_rows [l_ks, 2 (a, b), K*N] NTT form."""
        c = _u64(ct, (2, self.KN))
        k = _u64(key_rows, (self.scheme["l_ks"], 2, self.KN))
        out = np.zeros((2, self.KN), dtype=np.uint64)
        self.check(self.lib.opir_bv_ks(self.h, _ptr(c), int(galois_k), _ptr(k), _ptr(out)))
        return out

    def mod_switch(self, ct):
        c = _u64(ct, (2, self.KN))
        out = np.zeros((2, self.N), dtype=np.uint64)
        self.check(self.lib.opir_mod_switch(self.h, _ptr(c), _ptr(out)))
        return out


class GpuDatabase:
    """One opir_db: a resident NTT-form database of one layout."""

    def __init__(self, ctx: GpuContext, fst_dim_sz, other_dim_sz, num_dims, expansion_height, _view_of=None):
        self.ctx = ctx
        self.lib = ctx.lib
        self.N = ctx.N
        self.KN = ctx.KN
        self.N0, self.Nrest, self.dims, self.h = fst_dim_sz, other_dim_sz, num_dims, expansion_height
        self.u = fst_dim_sz + ctx.scheme["l_ep"] * (num_dims - 1)
        h = ctypes.c_void_p()
        if _view_of is None:
            ctx.check(self.lib.opir_db_create(ctx.h, fst_dim_sz, other_dim_sz, num_dims, expansion_height,
                                              ctypes.byref(h)))
        else:
            ctx.check(self.lib.opir_db_view(_view_of.h, ctx.h, ctypes.byref(h)))
        self.h = h
        self._view_of = _view_of  # a view keeps the database it shares alive
        ctx._dbs.add(self)

    def view(self, ctx: "GpuContext"):
        """A handle on `ctx` (a clone of this database's context) sharing this
        resident database, with its own workspace (opir_db_view)."""
        return GpuDatabase(ctx, self.N0, self.Nrest, self.dims, self.h, _view_of=self)

    def close(self):
        if getattr(self, "h", None):
            self.lib.opir_db_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_plain(self, pts, first_pt=0):
        pts = np.asarray(pts)
        if pts.size % self.N:
            raise InvalidArgument("plaintexts must hold a whole number of N-coefficient polynomials")
        # server.cpp:129-137, checked before any narrowing so no value wraps
        if pts.size and (int(pts.max()) >= self.ctx.scheme["plain_mod"] or int(pts.min()) < 0):
            raise InvalidArgument("plaintext coefficient not reduced mod t")
        if self.ctx.scheme["plain_mod"] <= 1 << 16:
            a = np.ascontiguousarray(pts, dtype=np.uint16)
            self.ctx.check(self.lib.opir_db_load_plain(self.h, _ptr(a), first_pt, a.size // self.N))
        else:
            a = _u64(pts)
            self.ctx.check(self.lib.opir_db_load_plain_u64(self.h, _ptr(a), first_pt, a.size // self.N))

    def load_ntt(self, lo, hi):
        n = self.N * self.N0 * self.Nrest
        lo, hi = np.asarray(lo), np.asarray(hi)
        if lo.size != n or hi.size != n:
            raise InvalidArgument(f"reference-layout database needs {n} residues per limb")
        if int(lo.max()) >= 1 << 32 or int(hi.max()) >= 1 << 32:
            raise InvalidArgument("database residue does not fit 32 bits")
        lo = np.ascontiguousarray(lo, dtype=np.uint32)
        hi = np.ascontiguousarray(hi, dtype=np.uint32)
        self.ctx.check(self.lib.opir_db_load_ntt(self.h, _ptr(lo), _ptr(hi)))

    def answer(self, client_id, query):
        q = _u64(query, (2, self.KN))
        out = np.zeros((2, self.N), dtype=np.uint64)
        self.ctx.check(self.lib.opir_answer(self.h, client_id, _ptr(q), _ptr(out)))
        return out

    def answer_batch(self, client_ids, queries):
        ids = np.ascontiguousarray(client_ids, dtype=np.uint64)
        q = _u64(queries, (len(ids), 2, self.KN))
        out = np.zeros((len(ids), 2, self.N), dtype=np.uint64)
        self.ctx.check(self.lib.opir_answer_batch(self.h, len(ids), _ptr(ids), _ptr(q), _ptr(out)))
        return out

    def answer_batch_device(self, client_ids, d_queries_ptr, d_resps_ptr):
        ids = np.ascontiguousarray(client_ids, dtype=np.uint64)
        self.ctx.check(self.lib.opir_answer_batch_device(self.h, len(ids), _ptr(ids),
                                                         ctypes.c_void_p(d_queries_ptr),
                                                         ctypes.c_void_p(d_resps_ptr)))

    def query_gen(self, keys, counters, indices):
        """Queries of seeded clients generated on the device: [B, 2, K*N]."""
        kb = np.frombuffer(b"".join(bytes(k) for k in keys), dtype=np.uint8).copy()
        c = np.ascontiguousarray(counters, dtype=np.uint64)
        i = np.ascontiguousarray(indices, dtype=np.uint64)
        if kb.size != 32 * len(c) or len(c) != len(i):
            raise InvalidArgument("one key, counter and index per query")
        out = np.zeros((len(c), 2, self.KN), np.uint64)
        self.ctx.check(self.lib.opir_query_gen(self.h, len(c), _ptr(kb), _ptr(c), _ptr(i), _ptr(out)))
        return out

    def reserve(self, batch):
        """Prepare answers of up to `batch` queries: workspace, host staging
        and, in autotune mode, the scan shape of every batch-size class."""
        self.ctx.check(self.lib.opir_db_reserve(self.h, batch))

    def trim(self, batch):
        """Give back the workspace beyond `batch` queries (0: all of it)."""
        self.ctx.check(self.lib.opir_db_trim(self.h, batch))

    def workspace_bytes(self, batch):
        """Device bytes of the answer workspace for `batch` queries."""
        out = ctypes.c_size_t()
        self.ctx.check(self.lib.opir_db_workspace_bytes(self.h, batch, ctypes.byref(out)))
        return out.value

    def tuned_scan_variant(self, batch):
        """Scan shape the autotuner picked for this batch size (-1: not tuned)."""
        out = ctypes.c_int(-1)
        self.ctx.check(self.lib.opir_db_tuned_scan_variant(self.h, batch, ctypes.byref(out)))
        return out.value

    def expand(self, client_id, query):
        q = _u64(query, (2, self.KN))
        out = np.zeros((self.u, 2, self.KN), dtype=np.uint64)
        self.ctx.check(self.lib.opir_expand(self.h, client_id, _ptr(q), _ptr(out)))
        return out

    def complete(self, client_id, expanded):
        e = _u64(expanded, (self.u, 2, self.KN))
        ns = self.dims - 1
        out = np.zeros((ns, 2 * self.ctx.scheme["l_ep"], 2 * self.KN), dtype=np.uint64)
        self.ctx.check(self.lib.opir_complete(self.h, client_id, _ptr(e), _ptr(out)))
        return out

    def first_dim(self, query_cts):
        q = _u64(query_cts, (self.N0, 2, self.KN))
        out = np.zeros((self.Nrest, 2, self.KN), dtype=np.uint64)
        self.ctx.check(self.lib.opir_first_dim(self.h, _ptr(q), _ptr(out)))
        return out

    def other_dim(self, mid, sel):
        m = _u64(mid, (self.Nrest, 2, self.KN))
        ns = self.dims - 1
        s = _u64(sel, (ns, 2 * self.ctx.scheme["l_ep"], 2 * self.KN)) if ns else np.zeros(1, dtype=np.uint64)
        out = np.zeros((2, self.KN), dtype=np.uint64)
        self.ctx.check(self.lib.opir_other_dim(self.h, _ptr(m), _ptr(s), _ptr(out)))
        return out


class GpuTree:
    """One opir_tree: the per-level packed plaintexts of a tree PIR database."""

    def __init__(self, ctx: GpuContext, L, a):
        self.ctx = ctx
        self.lib = ctx.lib
        self.L, self.a = L, a
        h = ctypes.c_void_p()
        ctx.check(self.lib.opir_tree_create(ctx.h, L, a, ctypes.byref(h)))
        self.h = h
        vals = [ctypes.c_size_t() for _ in range(4)]
        ctx.check(self.lib.opir_tree_shape(h, *[ctypes.byref(v) for v in vals]))
        self.w, self.h_q, self.N0, self.nsel = (v.value for v in vals)
        ctx._dbs.add(self)

    def close(self):
        if getattr(self, "h", None):
            self.lib.opir_tree_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_level(self, level, plaintexts):
        a = _u64(plaintexts)
        if a.size % self.ctx.N:
            raise InvalidArgument("plaintexts must hold whole polynomials")
        self.ctx.check(self.lib.opir_tree_load_level(self.h, level, _ptr(a)))

    def select(self, client_id, query):
        """Every level's selected ciphertext, [L + 1, 2, K*N], coefficient form."""
        q = _u64(query, (2, self.ctx.KN))
        out = np.zeros((self.L + 1, 2, self.ctx.KN), np.uint64)
        self.ctx.check(self.lib.opir_tree_select(self.h, client_id, _ptr(q), _ptr(out)))
        return out


def answer_batch_device_raw(db_handle, client_ids, d_queries_ptr, d_resps_ptr):
    """opir_answer_batch_device on a database handle owned by the host API."""
    lib = load_library()
    ids = np.ascontiguousarray(client_ids, dtype=np.uint64)
    rc = lib.opir_answer_batch_device(ctypes.c_void_p(db_handle), len(ids), _ptr(ids),
                                      ctypes.c_void_p(d_queries_ptr), ctypes.c_void_p(d_resps_ptr))
    return rc


def ctx_call_raw(name, ctx_handle, *args):
    lib = load_library()
    return getattr(lib, name)(ctypes.c_void_p(ctx_handle), *args)


def last_error_raw(ctx_handle):
    return load_library().opir_last_error(ctypes.c_void_p(ctx_handle)).decode()
