/*
 * onionpir_gpu.h -- C ABI of the MI355X (gfx950) OnionPIRv2 server library.
 *
 * The host-side PirServer (onionpirv2_amd/host/onionpir_host.h) calls only these
 * entry points; any other host language binds them the same way (ctypes,
 * cgo, JNI). Plain pointers and sizes only. Every call returns an int status
 * (OPIR_OK = 0, negative on error); opir_last_error() gives the message.
 * The status maps onto the reference's exceptions:
 *   OPIR_EINVAL -> std::invalid_argument, OPIR_ERANGE -> std::out_of_range,
 *   others      -> std::runtime_error.
 *
 * Data formats at this boundary are the reference's own, with K RNS limbs
 * (K = 1 logical limb mod Q = q1*q2 for CONFIG_N2048_K1_COMP, K = 1 limb mod
 * one ~60-bit prime for CONFIG_N2048_K1, K = 2 limbs mod q1 and q2 for
 * CONFIG_N2048_K2_MP and CONFIG_N4096_K2_MP):
 *   - ciphertext: c0 then c1, each K*N uint64 values, limb-major
 *     ([limb 0 N][limb 1 N] when K = 2), src/includes/rlwe.h:17-31;
 *   - NTT-form values use the reference's NTT evaluation order;
 *   - RGSW rows: 2*l rows of [c0 K*N][c1 K*N], src/includes/gsw.h:8-17;
 *   - BV keys: [h keys][L_KS rows][a, b][K*N], galois k = N/2^i + 1,
 *     src/includes/bv_keyswitch.h:45-66;
 *   - response: c0 then c1, N uint64 values mod small_q (one limb after the
 *     mod switch);
 *   - the stage entry points below take the same K*N-value polynomials.
 * Host buffers are owned by the caller; device buffers by the context.
 * A context is not thread-safe (the reference's PirServer is not either,
 * src/includes/gsw.h:25-30); one context serves one host thread.
 */
#ifndef ONIONPIR_GPU_H
#define ONIONPIR_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OPIR_OK 0
#define OPIR_EINVAL (-1)
#define OPIR_ERANGE (-2)
#define OPIR_ERUNTIME (-3)
#define OPIR_EHIP (-4)
#define OPIR_ENOMEM (-5)

typedef struct opir_ctx opir_ctx;
typedef struct opir_db opir_db;

/* Scheme description: the reference's compile-time DBConsts plus the
 * PirParams-derived values (src/includes/database_constants.h:77-94,
 * src/pir.cpp:14-98), as a runtime value. */
typedef struct opir_scheme {
  uint32_t poly_degree;     /* N: 2048 or 4096 (CONFIG_N4096_K2_MP) */
  uint32_t composite;       /* 1: Q = q1*q2 composite single limb (CONFIG_N2048_K1_COMP);
                               0: the RNS limbs q1, q2 of the ciphertext (CONFIG_N2048_K2_MP,
                               CONFIG_N4096_K2_MP, database_constants.h:96-124), or one prime q1
                               when q2 = 0 (CONFIG_N2048_K1, database_constants.h:61-75) */
  uint64_t q1, q2;          /* RNS primes of Q (q2 = 0: one prime); at N = 2048 with two
                               primes (CONFIG_N2048_K1_COMP, CONFIG_N2048_K2_MP) each is
                               below 2^29, as the reference's 29-bit primes are */
  uint64_t ntt_root;        /* composite: primitive 2N-th root mod Q (w_crt); otherwise root mod q1 */
  uint64_t plain_mod;       /* t */
  uint64_t small_q;         /* response modulus q' */
  uint32_t l_ep, base_log2;         /* data selector gadget */
  uint32_t l_key, base_log2_key;    /* RGSW(s) completion gadget */
  uint32_t l_ks, bv_base_log2;      /* BV keyswitch gadget */
  uint64_t ntt_root2;       /* two primes: primitive 2N-th root mod q2 (ignored otherwise) */
  double noise_std_dev;     /* error standard deviation of device-generated keys and queries
                               (DBConsts::NoiseStdDev, database_constants.h:48); 0 -> 2.55 */
} opir_scheme;

/* Replaces the scheme set-up of PirParams::PirParams / init_composite_rns
 * (src/pir.cpp:14-98). device = HIP device ordinal. */
int opir_ctx_create(const opir_scheme *scheme, int device, opir_ctx **out);
void opir_ctx_destroy(opir_ctx *ctx);
/* Message of the last failed call on ctx (or of the last failed
 * opir_ctx_create when ctx is NULL). Valid until the next call. */
const char *opir_last_error(const opir_ctx *ctx);
/* Use an existing HIP stream (hipStream_t) for all work of this context;
 * NULL reverts to the context's own stream. */
int opir_ctx_set_stream(opir_ctx *ctx, void *hip_stream);
/* Restrict the context's own streams to a set of compute units (bit i of
 * mask word i/32 = CU i, as hipExtStreamCreateWithCUMask takes it; words = 0
 * restores unrestricted streams). For a server that shares the device
 * between databases: a context whose answers stream a very large database
 * can leave some CUs to the others, so their short kernels need not wait for
 * its long-running workgroups. No counterpart in the reference (one CPU
 * server per process). Masked streams follow the default stream semantics
 * of hipExtStreamCreateWithCUMask. */
int opir_ctx_set_cu_mask(opir_ctx *ctx, const uint32_t *mask, size_t words);
/* Serving support, no counterpart in the reference: a clone is a context on
 * the same device with the parent's scheme tables and client keys (read
 * through the parent; key setters on a clone fail with OPIR_EINVAL) but its
 * own streams, events and staging, so answers issued on a clone and on its
 * parent from different host threads run concurrently on the device. A view
 * is a database handle on such a context that shares src's resident
 * database (no copy) with its own answer workspace. The parent context (and
 * src) must outlive its clones (views); keys must not change while clones
 * answer. */
int opir_ctx_clone(opir_ctx *parent, opir_ctx **out);
int opir_db_view(opir_db *src, opir_ctx *ctx, opir_db **out);
/* Blocks until all work queued by the context has finished. */
int opir_ctx_synchronize(opir_ctx *ctx);

/* Replaces PirServer::PirServer(const PirParams&) database allocation
 * (src/server.cpp:29-46) for a layout PirParams::with_layout produced
 * (src/pir.cpp:100-134): N0 = fst_dim_sz, Nrest = other_dim_sz. */
int opir_db_create(opir_ctx *ctx, size_t fst_dim_sz, size_t other_dim_sz, size_t num_dims,
                   size_t expansion_height, opir_db **out);
/* A database uses its context's stream and tables: destroy it before the
 * context. */
void opir_db_destroy(opir_db *db);
/* Replaces PirServer::load_data's NTT + transpose (src/server.cpp:118-171):
 * `count` coefficient-form plaintexts (N values < t each, pt-major),
 * logical indices [first_pt, first_pt + count). A coefficient >= t or a
 * range outside the layout -> OPIR_EINVAL (the reference throws
 * std::invalid_argument, src/server.cpp:129-137); nothing is loaded then. */
int opir_db_load_plain(opir_db *db, const uint16_t *coeffs, size_t first_pt, size_t count);
/* Loads an already preprocessed database in the reference's composite
 * layout: lo[coeff * num_pt + pt] = NTT_Q(pt)[coeff] mod q1, hi the same
 * mod q2 (src/server.cpp:162-170). A residue >= its limb prime ->
 * OPIR_EINVAL. */
int opir_db_load_ntt(opir_db *db, const uint32_t *lo, const uint32_t *hi);
/* Same as opir_db_load_plain with 64-bit coefficients (< t), for plaintext
 * moduli beyond 16 bits (CONFIG_N4096_K2_MP: t ~ 2^39). */
int opir_db_load_plain_u64(opir_db *db, const uint64_t *coeffs, size_t first_pt, size_t count);
/* Replaces PirServer::gen_data (src/server.cpp:53-66) for large databases:
 * fills every plaintext on the device, coefficient c of plaintext i =
 * splitmix64(seed + i*N + c) mod t (splitmix64 as in Vigna's reference
 * code), so no plaintext crosses PCIe. The host recomputes any recorded
 * plaintext with the same formula. */
int opir_db_gen_random(opir_db *db, uint64_t seed);

/* Replaces PirServer::set_client_bv_galois_key + set_client_gsw_key /
 * set_client_session_keys (src/server.cpp:857-875). bv: [h][l_ks][2: a, b][N]
 * mod Q in NTT form, h >= expansion height of every db used with it;
 * gsw: [2*l_key][2][N] mod Q in NTT form. Re-setting an id replaces the keys. */
int opir_client_set_keys(opir_ctx *ctx, uint64_t client_id, const uint64_t *bv, size_t h, const uint64_t *gsw);
/* Seed-compressed form of opir_client_set_keys (the reference models these
 * sizes, src/includes/pir.h:267-291, with no encoder): bv_b [h][l_ks][K*N]
 * holds the b half of every BV key row and gsw_c0 [2*l_key][K*N] the c0 half
 * of every RGSW(s) row, NTT form; the uniform halves (a, c1) are regenerated
 * on the device from the client's 32-byte key seed: row (level, r) of the BV
 * keys is uniform(seed, domain 4, level*l_ks + r, limb), RGSW(s) row r is
 * uniform(seed, domain 6, r, limb), with uniform(...) the ChaCha20 stream of
 * onionpirv2_amd/host/prg.h (coefficient k = the 128-bit draw k mod the
 * limb's modulus, Q for the composite scheme). Re-setting an id replaces
 * the keys. */
int opir_client_set_keys_seeded(opir_ctx *ctx, uint64_t client_id, const uint8_t *seed, const uint64_t *bv_b,
                                size_t h, const uint64_t *gsw_c0);
int opir_client_remove(opir_ctx *ctx, uint64_t client_id);
/* Reads a resident client's keys back in the layout opir_client_set_keys
 * takes (bv [h][l_ks][a, b][K*N], gsw [2*l_key][2][K*N], NTT form); h must
 * equal the resident BV key count. */
int opir_client_get_keys(opir_ctx *ctx, uint64_t client_id, uint64_t *bv, size_t h, uint64_t *gsw);

/* ---- Client-side helpers on the GPU (load generation) ----
 * No counterpart on the reference's server: these generate, in HBM, what a
 * seeded client (the host PirClient(params, id, key32), whose secret, uniform
 * halves and errors come from the ChaCha20 sampler of
 * onionpirv2_amd/host/prg.h) generates on its CPU -- the same values -- for
 * PirClient::create_bv_galois_keys / generate_gsw_from_key
 * (src/bv_keyswitch.cpp:191-276, src/client.cpp:25-38) and
 * fast_generate_query (src/client.cpp:107-225). The 32-bit-limb schemes
 * (CONFIG_N2048_K1_COMP, CONFIG_N2048_K2_MP) only; a 64-bit-limb context
 * (CONFIG_N4096_K2_MP, CONFIG_N2048_K1) returns OPIR_EINVAL.
 * keygen: n clients (keys [n][32] bytes), h BV levels each, installed as
 * resident keys under ids[i] (no key material crosses PCIe).
 * query_gen: B queries [B][2][K*N] of clients keys[i] with query counter
 * counters[i] (the host client's fast_generate_query numbers its queries
 * from 0) for plaintext indices[i] of db's layout; _device writes device
 * memory, the other host memory. */
int opir_client_keygen_device(opir_ctx *ctx, size_t n, const uint64_t *ids, const uint8_t *keys, size_t h);
int opir_query_gen(opir_db *db, size_t B, const uint8_t *keys, const uint64_t *counters, const uint64_t *indices,
                   uint64_t *queries);
int opir_query_gen_device(opir_db *db, size_t B, const uint8_t *keys, const uint64_t *counters,
                          const uint64_t *indices, uint64_t *d_queries);

/* Replaces PirServer::make_query (src/server.cpp:895-947): query is the
 * coefficient-form packed BFV query [2][N] mod Q; resp receives the
 * mod-switched response [2][N] mod small_q. Unknown client -> OPIR_ERANGE. */
int opir_answer(opir_db *db, uint64_t client_id, const uint64_t *query, uint64_t *resp);
/* Batched form: B queries (distinct or repeated client ids) against one
 * database in one pass; queries [B][2][N], resps [B][2][N]. */
int opir_answer_batch(opir_db *db, size_t B, const uint64_t *client_ids, const uint64_t *queries,
                      uint64_t *resps);
/* Same as opir_answer_batch with queries/resps in device memory; work is
 * queued on the context stream and this call returns without waiting.
 * Client ids and keys are checked as in opir_answer_batch; the query values
 * are not (they are never read on the host): the caller guarantees every
 * value is reduced mod Q (per limb for the two-prime schemes). An unreduced
 * value yields a wrong response, never an access outside the buffers. */
int opir_answer_batch_device(opir_db *db, size_t B, const uint64_t *client_ids, const uint64_t *d_queries,
                             uint64_t *d_resps);

/* Serving support. No counterpart in the reference, whose PirServer answers
 * one query per make_query call; a serving loop uses these to admit batches
 * by HBM budget. workspace_bytes: device bytes of the answer workspace for
 * batches of up to B queries against db. reserve: allocate that workspace
 * now (it is kept and only grows), with the pinned host staging of
 * opir_answer_batch for B queries, and in autotune mode settle the scan tile
 * shape of every batch-size class up to B; answers of up to B queries then
 * neither allocate nor stop to time shapes. key_bytes: device bytes of one
 * client's keys at expansion height h. */
int opir_db_workspace_bytes(const opir_db *db, size_t B, size_t *out);
int opir_db_reserve(opir_db *db, size_t B);
/* Gives back the part of db's answer workspace beyond B queries (it is
 * reallocated for B; B = 0 frees it), after the work queued on the context
 * has finished; a later answer of more queries allocates again. A serving
 * loop that re-plans its batch limits returns an idle database's extra
 * workspace to the budget this way. */
int opir_db_trim(opir_db *db, size_t B);
int opir_client_key_bytes(const opir_ctx *ctx, size_t h, size_t *out);
/* Device bytes of all client keys resident on ctx now. */
int opir_ctx_resident_key_bytes(const opir_ctx *ctx, size_t *out);

/* Per-stage device time of the last answer call under the reference's stage
 * names (src/includes/logging.h:122-134): [0] Expand, [1] Convert,
 * [2] First dim, [3] Other dim, [4] Modulus switching, [5] scan kernel only.
 * Convert (selector completion) runs on a second stream concurrently with
 * First dim; both are timed from the end of Expand. For N = 4096 two-prime
 * batches whose completion has a workgroup per compute unit or more, it runs
 * ahead of the scan on the answer stream instead, and First dim includes it.
 * Requires opir_ctx_set_profiling(ctx, 1) before the answer call. */
int opir_ctx_set_profiling(opir_ctx *ctx, int on);
int opir_last_stage_ms(opir_ctx *ctx, float *out6);

/* First-dimension scan tile shape (no counterpart in the reference, whose
 * scan has one CPU loop nest): OPIR_SCAN_DEFAULT (-1, the default) picks the
 * shape per batch size from a table measured on MI355X; OPIR_SCAN_AUTOTUNE
 * (-2) times every applicable shape on the live operands the first time a
 * database sees a batch-size class (powers of two of 2B) and keeps the
 * fastest; 0 .. count-1 forces one, for tuning and for tests that check
 * every shape against the same answers. A forced shape that cannot take a
 * batch size makes the answer call fail with OPIR_EINVAL. Results are
 * identical for every shape. For the 29-bit limbs the shapes include a VALU
 * form (u32 x u32 -> u64 multiply-accumulate) beside the int8 MFMA ones, so
 * autotuning measures where the matrix cores pay off. */
#define OPIR_SCAN_DEFAULT (-1)
#define OPIR_SCAN_AUTOTUNE (-2)
int opir_ctx_set_scan_variant(opir_ctx *ctx, int variant);
int opir_scan_variant_count(const opir_ctx *ctx, int *out);
/* The shape autotuning picked for B queries on db (-1 if not tuned yet). */
int opir_db_tuned_scan_variant(const opir_db *db, size_t B, int *out);
/* Launch mode of the answer calls that replace PirServer::make_query
 * (src/server.cpp:895-947): HIP graphs for the batched form's
 * per-dimension launches. on != 0: the first answer for a given batch size, key-pointer
 * table slot, query/response buffers and scan shape captures its whole
 * launch sequence (expansion, completion on the auxiliary stream, scan, MUX
 * levels, mod switch) into a graph; later calls with the same ones replay
 * it (up to 16 graphs per database, least recently used dropped). Answers
 * are identical either way. Profiling and a scan shape still being
 * autotuned run without a graph. Off by default. */
int opir_ctx_set_graphs(opir_ctx *ctx, int on);

/* ---- Tree PIR on the same kernels (32-bit-limb schemes) ----
 * The server side of the reference's full-binary-tree PIR layer for a tree of
 * 2^L leaves with first dimension N0 = 2^a (TreePirParams from
 * make_tree_pir_params_for_scheme, src/tree_query.cpp:44-61, ell_beta =
 * ell_gamma = l_ep; L >= 11 + a at N = 2048).
 * load_level: level l's R_l packed plaintexts [R_l][N] (< t), the layout of
 *   pack_tree_level (src/tree_select.cpp:52-74), R_l = 2^max(l - 11, 0).
 * select: one packed tree query [2][K*N] (make_tree_query,
 *   src/tree_query.cpp:140-196) of a client whose keys reach expansion height
 *   h_q; out receives, for every level l = 0..L, select_level's ciphertext
 *   [2][K*N] in coefficient form (unpack_tree_query + build_alpha_pyramid +
 *   select_level, src/tree_query.cpp:198-240, src/tree_select.cpp:121-238).
 * shape: the packed width w, expansion height h_q, N0 and selector count
 *   b + r of the tree's query. */
typedef struct opir_tree opir_tree;
int opir_tree_create(opir_ctx *ctx, size_t L, size_t a, opir_tree **out);
void opir_tree_destroy(opir_tree *tree);
int opir_tree_shape(const opir_tree *tree, size_t *w, size_t *h_q, size_t *N0, size_t *nsel);
int opir_tree_load_level(opir_tree *tree, size_t level, const uint64_t *coeffs);
int opir_tree_select(opir_tree *tree, uint64_t client_id, const uint64_t *query, uint64_t *out);

/* ---- Stage entry points (host buffers), one per reference stage ---- */

/* utils::ntt_fwd / ntt_inv over Q (src/utils.cpp:84-90), in place on
 * `count` polynomials of N values. */
int opir_ntt_fwd(opir_ctx *ctx, uint64_t *data, size_t count);
int opir_ntt_inv(opir_ctx *ctx, uint64_t *data, size_t count);
/* PirServer::fast_expand_qry (src/server.cpp:731-816): out receives the
 * u = N0 + l_ep*(dims-1) leaves, [u][2][N] coefficient form. */
int opir_expand(opir_db *db, uint64_t client_id, const uint64_t *query, uint64_t *out);
/* PirServer::complete_selectors (src/server.cpp:823-855): expanded holds
 * the u leaves; sel receives [dims-1][2*l_ep][2][N] NTT form. */
int opir_complete(opir_db *db, uint64_t client_id, const uint64_t *expanded, uint64_t *sel);
/* PirServer::evaluate_first_dim (src/server.cpp:313-361, 529-597): query
 * [N0][2][N] coefficient form; out [Nrest][2][N] coefficient form. */
int opir_first_dim(opir_db *db, const uint64_t *query, uint64_t *out);
/* PirServer::evaluate_other_dim (src/server.cpp:599-722): mid [Nrest][2][N]
 * coefficient form, sel [dims-1][2*l_ep][2][N] NTT form; out [2][N]
 * coefficient form. */
int opir_other_dim(opir_db *db, const uint64_t *mid, const uint64_t *sel, uint64_t *out);
/* GSWEval::external_product (src/gsw.cpp:119-182): gsw [2l][2][N] NTT form,
 * ct [2][N] coefficient form -> out [2][N] NTT form. */
int opir_ext_prod(opir_ctx *ctx, const uint64_t *gsw, size_t l, size_t base_log2, const uint64_t *ct,
                  uint64_t *out);
/* PirServer::ext_prod_mux (src/server.cpp:652-722, declared in
 * src/includes/server.h:96-99): out = x + INTT(gsw [x] (y - x)), the
 * homomorphic select of y (gsw encrypts 1) or x (gsw encrypts 0). gsw
 * [2l][2][N] NTT form; x, y, out [2][N] coefficient form. x and y are left
 * unchanged (the reference uses y as scratch; its contents after the call
 * are not part of the result). */
int opir_ext_prod_mux(opir_ctx *ctx, const uint64_t *gsw, size_t l, size_t base_log2, const uint64_t *x,
                      const uint64_t *y, uint64_t *out);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the homomorphic
 * select of opir_ext_prod_mux with every operand in device memory (d_gsw
 * [2l][2][K*N] NTT form; d_x, d_y, d_out [2][K*N] coefficient form, values
 * reduced; d_out may alias d_x). It calls no allocator: its intermediates use
 * a fixed 2 MiB device array per GPU, so l is limited to 16 rows (32-bit-limb
 * schemes) or 32 (64-bit-limb schemes), larger -> OPIR_EINVAL. The work is queued on
 * the context stream and the call returns without waiting. All calls on one
 * GPU share that array: calls that may overlap on the device (other streams,
 * other contexts on the same GPU) race. */
int opir_ext_prod_mux_device_noalloc(opir_ctx *ctx, const uint64_t *d_gsw, size_t l, size_t base_log2,
                                     const uint64_t *d_x, const uint64_t *d_y, uint64_t *d_out);
/* bvks::bv_apply_galois_inplace (src/bv_keyswitch.cpp:296-367 for K = 1,
 * 373-481 for K = 2): ct [2][K*N] coefficient form; key: the l_ks rows of one
 * BvKeySwitchKey for galois element k, [l_ks][a, b][K*N] NTT form
 * (src/includes/bv_keyswitch.h:48-60); out [2][K*N] coefficient form, the
 * keyswitched sigma_k(ct) under the original secret. k must be odd and < 2N. */
int opir_bv_ks(opir_ctx *ctx, const uint64_t *ct, uint32_t galois_k, const uint64_t *key, uint64_t *out);
/* PirServer::mod_switch_inplace (src/server.cpp:1025-1043): ct [2][N]
 * coefficient form mod Q -> out [2][N] mod small_q. */
int opir_mod_switch(opir_ctx *ctx, const uint64_t *ct, uint64_t *out);

#ifdef __cplusplus
}
#endif

#endif /* ONIONPIR_GPU_H */
