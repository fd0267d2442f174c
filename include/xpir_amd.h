/*
 * xpir_amd.h -- C ABI of the MI355X NFLLWE PIR reply engine (libxpir_amd.so).
 *
 * One shared library, plain pointers and sizes, integer status codes, opaque handles.
 * Buffers are host memory unless the name ends in _dev (device memory on the
 * context's GPU).  Word layouts are the reference's:
 *   poly         m blocks of n uint64 (crypto/NFLlib.hpp:72, NFLLWEDatatypes.hpp:20-36)
 *   ciphertext   a[m][n] || b[m][n], 16*n*m bytes (NFLLWEPublicParameters.cpp:148-151)
 *   query level  dims[l] ciphertexts, canonical words (NFLLWE.cpp:355-356)
 *   NTT DB       units x P x m x n lazy words in [0,2p) (NFLlib.cpp:375-401)
 *   reply        P_d ciphertexts, lazy words (PIRReplyGeneratorNFL_internal.cpp:570-574)
 * Every reply is bit-identical to the reference's generateReply on the same inputs.
 *
 * Threading: one in-flight call per context; contexts are independent.  The serving
 * loop (xpg_server_*) owns its own concurrency.
 */
#ifndef XPIR_AMD_H
#define XPIR_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference prints or exit()s; this ABI returns) ---- */
#define XPG_OK 0
#define XPG_EINVAL -1      /* malformed argument / unsupported parameters */
#define XPG_ENOMEM -2      /* device or host allocation failed             */
#define XPG_EDEVICE -3     /* HIP runtime error                            */
#define XPG_ESIZE -4       /* buffer too small / size mismatch             */
#define XPG_EBUSY -5       /* admission rejected (serving loop full)       */

#define XPG_MAX_LEVELS 10  /* PIRParameters.n[MAX_REC_LVL], pir/PIRParameters.hpp:22 */

typedef struct xpg_ctx xpg_ctx;
typedef struct xpg_db xpg_db;

/* Cryptosystem parameters for one "LWE:k:n:bits" string and first dimension n0.
 * Replaces HomomorphicCryptoFactory::getCryptoMethod (pir/libpir.cpp:20-23) ->
 * NFLLWE::setNewParameters (crypto/NFLLWE.cpp:55-120) and
 * setandgetAbsBitPerCiphertext(n0) (crypto/NFLLWE.cpp:656-668). */
typedef struct {
  uint32_t k, n, logn, m, bits, Berr;
  int32_t A;                  /* absorption bits per coefficient                 */
  uint64_t n0;
  uint64_t moduli[4];
  uint64_t abs_bytes;         /* A*n/8: plaintext bytes per poly (getAbsorptionBitsize/8) */
  uint64_t ct_bytes;          /* 16*n*m: ciphertext bytes (getCiphertextBitsize/8)        */
} xpg_params;

int xpg_params_from_string(const char *desc, uint64_t n0, xpg_params *out);

/* polys per import unit of `unit_bytes` bytes: ceil(8*unit_bytes/(A*n)) (NFLlib.cpp:378) */
uint32_t xpg_polys_for_bytes(const xpg_params *prm, uint64_t unit_bytes);
/* reply ciphertexts for d levels: computeReplySizeInChunks (PIRReplyGeneratorNFL_internal.cpp:667-679) */
uint32_t xpg_reply_polys(const xpg_params *prm, uint32_t P0, uint32_t d);

/* ---- context ---- */
int xpg_ctx_create(int device, xpg_ctx **out);
void xpg_ctx_destroy(xpg_ctx *ctx);
/* message of the last failure on this context ("" if none) */
const char *xpg_last_error(const xpg_ctx *ctx);
/* the context's HIP stream (hipStream_t); every call below is ordered on it */
void *xpg_ctx_stream(xpg_ctx *ctx);
int xpg_sync(xpg_ctx *ctx);

/* tuning knobs: "mac_kernel" = 0 (by shape), 1 (one chain per lane), 3 (streaming block
 * summaries; what 0 picks for few, long chains); "stream_e" = elements per streaming
 * block (0 = sized for one pass), "stream_u" = elements per load chunk 4|8|16 (0 = 8),
 * "stream_pf" = the chunks' load schedule: 0 by rule (rolling), 1 double-buffered, 2
 * rolling (the next chunk's loads go into the registers each group of four elements
 * frees), 3 a chunk's loads then its arithmetic; "graphs" = 1 to replay each reply's
 * launches (every level, the recursion between them) from a HIP graph captured on the
 * second call with the same database, query buffer and reply buffer (the first runs
 * eagerly and sizes the context buffers; any option change or buffer growth re-captures;
 * with timing on, replies run eagerly).  No reference counterpart; every setting yields
 * the same reply bytes.  "group_budget" = device bytes a batch's levels 1..d-1 may take
 * at once (requests are grouped under it; 0 = 24 GiB).  "alpha_zero_pad" = 1: a partial
 * last alpha aggregate holds its real elements and zeros (the default 0 gives it the
 * reference's bytes, see xpg_db_import); this one changes reply bytes when n_elems is
 * not a multiple of alpha.
 * "db_layout" = 1: databases imported (or loaded) from then on keep level 0 in the matrix
 * layout, for the int8 matrix cores: each word as eight signed byte digits in the
 * fragments of v_mfma_i32_32x32x32_i8, plus each group's last 40 elements as NTT words;
 * the row-major words are freed after the conversion (both exist during it).  Replies
 * on such a database take level 0 as the exact residue of all but the last 40 products
 * of each chain (matrix cores), then the reference's step for the last 40 from that
 * residue; a chain whose last steps cannot settle its excess multiple of p, and every
 * chain of a query holding a word not below its modulus, is recomputed whole with the
 * reference's step: the reply bytes are the row-major layout's.  Applies when
 * 72 <= dims[0] < 2^27 and m*n is a multiple of 256 (other databases keep the row-major
 * words); the conversion fails with XPG_EINVAL for a word >= 2^63 - 2^56 (none that an
 * import makes).  Such a database has no row-major words: xpg_db_export_ntt_words,
 * xpg_db_save and xpg_autotune refuse it (XPG_EINVAL); xpg_reply*, xpg_reply_batch_dev,
 * xpg_reply_dev_noalloc (when its scratch holds the launch's query fragments) and the
 * serving loop answer from it.  No reference counterpart. */
int xpg_ctx_set_option(xpg_ctx *ctx, const char *key, int64_t value);

/* device memory plumbing for callers without a GPU runtime of their own */
int xpg_dev_alloc(xpg_ctx *ctx, uint64_t bytes, void **ptr_dev);
int xpg_dev_free(xpg_ctx *ctx, void *ptr_dev);
int xpg_h2d(xpg_ctx *ctx, void *dst_dev, const void *src, uint64_t bytes);
int xpg_d2h(xpg_ctx *ctx, void *dst, const void *src_dev, uint64_t bytes);
/* page-locked host memory (hipHostMalloc) and a copy from it that returns at once,
 * ordered on the context stream (xpg_sync waits for it): staging for callers that
 * upload a query while it arrives.  No reference counterpart (pushQuery keeps the
 * elements in host memory, PIRReplyGeneratorNFL_internal.cpp:738-781). */
int xpg_host_alloc(xpg_ctx *ctx, uint64_t bytes, void **ptr);
int xpg_host_free(xpg_ctx *ctx, void *ptr);
int xpg_h2d_async(xpg_ctx *ctx, void *dst_dev, const void *src_pinned, uint64_t bytes);

/* ---- database import (PIRReplyGenerator::importData, pir/libpir.cpp:90-103 ->
 *      PIRReplyGeneratorNFL_internal::importDataNFL, :79-158) ----
 * elems: n_elems elements of elem_bytes, back to back (shorter elements zero-padded by
 * the caller, as DBVectorProcessor::readStream does, DBVectorProcessor.cpp:88-106).
 * alpha elements form one import unit (DBHandler::readAggregatedStream,
 * DBHandler.cpp:3-26).  When n_elems mod alpha = r > 0 the last unit gets the bytes the
 * reference gives it (its buffer is shared by every unit, and the bzero at
 * DBHandler.cpp:21-24 starts at the last real element's slot): elements 0..r-2 of the
 * unit, zeros in slots r-1..alpha-2, and in slot alpha-1 the previous unit's last element
 * (zeros if there is none); option "alpha_zero_pad" keeps the r elements and zeros the
 * rest instead.  dims[0..d-1] is the query grid; prod(dims) >= units; the grid's
 * missing units are zero polys (PIRReplyGeneratorNFL_internal.cpp:146-154).
 * The NTT-domain database stays resident in HBM until xpg_db_free, also past
 * xpg_ctx_destroy of the importing context, and serves replies on any context of the
 * same GPU (an imported_database outlives its PIRReplyGenerator in the reference's
 * keep-database sessions, apps/server/PIRServer.cpp:73-96). */
int xpg_db_import(xpg_ctx *ctx, const xpg_params *prm, const uint8_t *elems, uint64_t n_elems,
                  uint64_t elem_bytes, uint32_t alpha, uint32_t d, const uint32_t *dims,
                  xpg_db **out);
/* same, with the element bytes already in HBM (elems_dev) */
int xpg_db_import_dev(xpg_ctx *ctx, const xpg_params *prm, const uint8_t *elems_dev,
                      uint64_t n_elems, uint64_t elem_bytes, uint32_t alpha, uint32_t d,
                      const uint32_t *dims, xpg_db **out);
/* NTT-domain words produced elsewhere (units x P x m x n), e.g. by a CPU import */
int xpg_db_import_ntt_words(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *words,
                            uint64_t units, uint32_t P, uint64_t raw_elem_bytes, uint32_t d,
                            const uint32_t *dims, xpg_db **out);
/* copy the resident NTT words back (parity checks); cap in words */
int xpg_db_export_ntt_words(xpg_ctx *ctx, const xpg_db *db, uint64_t *out, uint64_t cap_words);
/* The NTT-domain database in a file, loaded straight back to HBM.  No reference
 * counterpart: the reference imports at every start (importDataNFL,
 * PIRReplyGeneratorNFL_internal.cpp:79-158) and keeps the imported database only in RAM
 * (apps/server/PIRServer.cpp:71-104).  The file is a 512-byte header (parameters, grid,
 * sizes, the partial-aggregate convention of the import, `digest`: 32 bytes the caller
 * computes over the raw database, or zeros when NULL; an order-keyed checksum of the
 * words) and the words as they lie in HBM; xpg_db_save writes it under a temporary name
 * of its own (path + ".tmp.<pid>.<n>") and renames it to path when complete.  xpg_db_load refuses (XPG_EINVAL) a file of another
 * format, made under other parameters than *prm, or with another digest (when digest is
 * not NULL), or whose words fail the checksum (recomputed on the device after the
 * copy); XPG_ESIZE for a file shorter or longer than its header says.  Host memory:
 * the context's two pinned staging chunks. */
int xpg_db_save(xpg_ctx *ctx, const xpg_db *db, const char *path, const uint8_t *digest);
int xpg_db_load(xpg_ctx *ctx, const xpg_params *prm, const char *path, const uint8_t *digest, xpg_db **out);
/* polys per unit at level 0, and the reply's ciphertext count */
uint32_t xpg_db_polys(const xpg_db *db);
uint32_t xpg_db_reply_polys(const xpg_db *db);
/* the recursion depth d, and dims[0..min(d, cap)-1] copied to dims (a loaded database's grid) */
uint32_t xpg_db_levels(const xpg_db *db, uint32_t *dims, uint32_t cap);
/* query words the database expects: sum_l dims[l] * 2*m*n */
uint64_t xpg_db_query_words(const xpg_db *db);
/* raw database bytes (n_elems * elem_bytes), the numerator of DB GB/s */
uint64_t xpg_db_raw_bytes(const xpg_db *db);
void xpg_db_free(xpg_db *db);

/* ---- reply generation (PIRReplyGenerator::generateReply, pir/libpir.cpp:106-135 ->
 *      PIRReplyGeneratorNFL_internal::generateReply, :340-445, with pushQuery's Shoup
 *      precompute (:738-781) folded into the kernels and popReply (libpir.cpp:144-154)
 *      replaced by one contiguous output buffer) ----
 * query: xpg_db_query_words(db) words, levels back to back; reply: P_d*2*m*n words.
 * Host buffers: the context keeps one device buffer of query + reply size for these
 * copies (grown on demand, released by xpg_ctx_destroy). */
int xpg_reply(xpg_ctx *ctx, const xpg_db *db, const uint64_t *query, uint64_t query_words,
              uint64_t *reply, uint64_t reply_cap_words);
/* device-resident query and reply, asynchronous on the context stream */
int xpg_reply_dev(xpg_ctx *ctx, const xpg_db *db, const uint64_t *query_dev, uint64_t *reply_dev);

/* xpg_reply_dev without any allocation, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No reference counterpart.  The same reply words, enqueued on the context stream; no
 * hipMalloc, hipFree, event or graph is created.  The intermediate-level buffers and the
 * streaming MAC's block summaries come from one static device array of
 * XPG_NOALLOC_SCRATCH_BYTES per GPU (held from library load until the process ends);
 * a reply that needs more returns XPG_ESIZE before anything is enqueued.  For d > 1 the
 * context must already hold the NTT tables of the database's parameters (any import or
 * ordinary reply on it builds them), else XPG_EINVAL.  Every call on one GPU shares
 * that GPU's array: calls on different contexts (streams) of the same GPU that overlap
 * in time race on it; calls on one context are ordered by its stream.  Nothing in the
 * array needs to start at zero (every word is written before it is read). */
#define XPG_NOALLOC_SCRATCH_BYTES (64ull << 20)
int xpg_reply_dev_noalloc(xpg_ctx *ctx, const xpg_db *db, const uint64_t *query_dev, uint64_t *reply_dev);

/* ---- batched replies: many queries over resident databases, the DB tiles of each
 *      database streamed from HBM once per query group.  No reference counterpart
 *      (the reference answers one query per generator); each reply is identical to
 *      xpg_reply_dev of the same query.  Requests are grouped per database in chunks of
 *      up to 64; level 0 of every chunk that runs register-blocked, over all the call's
 *      databases, is one persistent launch (several when their intermediates exceed
 *      "group_budget" or the launch's 24-segment table: a chunk takes one segment, two
 *      when its query count is above 4 and not a multiple of 4), followed by each
 *      chunk's recursion levels.  Option "batch_fuse" = 0 launches level 0 per chunk
 *      instead (same bytes).  Not for HIP graph capture: the launch uploads its tile
 *      table with a host-to-device copy (from the context's pinned table buffers). ---- */
typedef struct {
  const xpg_db *db;
  const uint64_t *query_dev;
  uint64_t *reply_dev;
} xpg_request;
int xpg_reply_batch_dev(xpg_ctx *ctx, const xpg_request *reqs, uint32_t n_reqs);

/* ---- serving loop: continuous batching over resident databases.  No reference
 *      counterpart (the reference serves one query per session thread,
 *      apps/server/PIRSession.cpp:371-493).  A dispatcher thread owns `workers` streams;
 *      pending requests are grouped per database and launched as one batch on the first
 *      idle worker, so arrivals during a batch overlap with it on the GPU.  The server's
 *      device memory is one pool of reply_pool_bytes, allocated at create: half holds the
 *      reply slots, the other half is split into one workspace per worker, from which a
 *      batch's intermediate levels are carved (nothing else is allocated per request).
 *      Admission: a request whose reply slot does not fit the free slots now gets
 *      XPG_EBUSY (submit again after releasing replies); one whose intermediates alone
 *      exceed a worker's workspace can never run and gets XPG_ESIZE.  A batch takes as
 *      many of one database's pending requests as a workspace holds (at most
 *      max_batch).  Batches are taken in arrival order of their oldest request, except
 *      that at most one heavy batch (more exact chain steps than heavy_steps) runs at a
 *      time: while one does, an idle worker takes the oldest light batch instead, so
 *      light requests never queue behind two heavy ones and heavy batches do not share
 *      the GPU with each other. ---- */
typedef struct xpg_server xpg_server;
typedef struct {
  uint64_t reply_pool_bytes;  /* the server's device memory: slots + workspaces */
  uint32_t workers;           /* batches in flight (one stream each)   */
  uint32_t max_batch;         /* queries per batch                     */
  uint64_t heavy_steps;       /* heavy batch threshold in exact chain steps (0 = 1e9,
                                 UINT64_MAX = every batch light)        */
} xpg_server_config;
typedef struct {
  uint64_t ticket;
  uint64_t *reply_dev;        /* valid until xpg_server_release(ticket) */
  uint64_t reply_words;
  double submit_us, start_us, done_us;  /* server clock (xpg_server_now_us) */
  int status;
} xpg_completion;
int xpg_server_create(int device, const xpg_server_config *cfg, xpg_server **out);
void xpg_server_destroy(xpg_server *srv);
/* query_dev stays owned by the caller and must stay valid until its completion */
int xpg_server_submit(xpg_server *srv, const xpg_db *db, const uint64_t *query_dev, uint64_t *ticket);
/* up to max completions; waits up to timeout_us when none is ready */
int xpg_server_poll(xpg_server *srv, xpg_completion *out, uint32_t max, uint32_t *n_out, int64_t timeout_us);
/* hands a completed ticket's reply slot back; XPG_EBUSY while the ticket is pending or
 * running (its slot may still be written), XPG_EINVAL for an unknown ticket */
int xpg_server_release(xpg_server *srv, uint64_t ticket);
double xpg_server_now_us(xpg_server *srv);
/* batches launched, requests in them, reply-slot bytes now held */
int xpg_server_stats(xpg_server *srv, uint64_t *batches, uint64_t *batched_requests, uint64_t *budget_used);

/* ---- kernel timing: HIP events around every MAC launch on the context stream ----
 * on = 1: around each MAC level and each level's main kernel; on = 2: around the main
 * kernels alone (fewer event records between launches: no gaps around a level's scan,
 * for timing the main kernel inside a throughput loop); 0: off */
int xpg_timing_enable(xpg_ctx *ctx, int on);
/* total milliseconds and launch count of the MAC kernels since the last reset */
int xpg_timing_read(xpg_ctx *ctx, double *mac_ms, uint64_t *mac_launches, int reset);
/* the same for the dominant kernel of each MAC alone (the streaming block-summary,
 * one-chain-per-lane or batched kernel, without its scan) */
int xpg_timing_read_main(xpg_ctx *ctx, double *ms, uint64_t *launches, int reset);
/* the same for the register-blocked kernel's launches alone (level 0 of d >= 2 replies
 * and of batches), with the exact chain steps they ran (queries x groups x P x m x n x 2
 * chains x n_l elements), for its integer-issue roofline */
int xpg_timing_read_batched(xpg_ctx *ctx, double *ms, uint64_t *launches, uint64_t *steps, int reset);

/* ---- MAC autotuning (no reference counterpart: the reference runs one CPU loop nest
 *      for every shape, pir/replyGenerator/PIRReplyGeneratorNFL_internal.cpp:511-595) ----
 * A MAC launch shape (one query's reply level, or a batch of queries on level 0) can run
 * on several kernels that all give the same words.  xpg_autotune times every candidate
 * on the levels of `db` (and on batches of 2, 4, ... max_batch queries on level 0),
 * checks each candidate's words against the untuned choice's (XPG_EDEVICE if any
 * differs), and records the fastest per shape in a table shared by every context on the
 * device; later launches of a recorded shape use it unless an option pins a kernel.
 * kind: 0 = one query's level, 1 = a batch (qn = batch size rounded up to a power of
 * two).  kernel: 1 = one chain per lane, 3 = streaming (a = chunk, b = double-buffered, c = elements per block,
 * 0 = one pass), 4 = register blocks (a rows x b queries per lane; c = 2: each wave
 * loads its query words, 3: a workgroup's four waves share them through LDS, 0: by shape;
 * built: 1x1, 1x2, 1x4, 4x1, 4x2, 4x4, 8x1, 8x2, 16x1 with c = 2 and 4x1, 4x2, 4x4, 8x1,
 * 8x2 with c = 3); a kind-1 entry
 * with kernel 0 records that the batch runs faster as single-query launches (the batch
 * call then takes that path for batches of this size on this database shape).  A
 * measured gain under 3 % keeps the fixed rule's choice. */
typedef struct {
  uint32_t kind, n, m, P;
  uint64_t n_l, groups;
  uint32_t qn;
  int32_t kernel, a, b, c;
  float ms;          /* measured ms per launch of this choice */
  float default_ms;  /* measured ms per launch of the untuned choice */
} xpg_tune_entry;
int xpg_autotune(xpg_ctx *ctx, const xpg_db *db, uint32_t max_batch, xpg_tune_entry *out, uint32_t cap,
                 uint32_t *count);
/* install entries (a saved table); XPG_EINVAL names a kernel or shape that is not built */
int xpg_tune_set(xpg_ctx *ctx, const xpg_tune_entry *entries, uint32_t count);
/* the device's table: up to cap entries into out, the total into *count */
int xpg_tune_get(xpg_ctx *ctx, xpg_tune_entry *out, uint32_t cap, uint32_t *count);
int xpg_tune_clear(xpg_ctx *ctx);

/* captured replies (option "graphs" = 1): replies captured into a HIP graph and replies
 * replayed from one since the context was created.  No reference counterpart. */
int xpg_graph_stats(xpg_ctx *ctx, uint64_t *captures, uint64_t *replays);

/* ---- client-side helpers (test and load-generation support; the reference's
 *      NFLLWE::enc/encrypt, crypto/NFLLWE.cpp:276-372,481-498, and
 *      PIRQueryGenerator_internal::generateQuery, :37-70) ---- */
/* secret key (m*n NTT-domain words) from a seed, uniform per modulus */
int xpg_keygen(const xpg_params *prm, uint64_t seed, uint64_t *sk);
/* decrypt n_ct ciphertexts (each a[m][n] || b[m][n], lazy words as replies hold them)
 * with the NTT-domain secret key sk_dev (m*n words) into n_ct * A*n/8 plaintext bytes
 * (plain_dev 8-byte aligned): NFLLWE::dec (crypto/NFLLWE.cpp:412-458) -> b - a*s,
 * invnttAndPowInvPhi (crypto/NFLlib.hpp:613-626), the centred CRT lift of poly2mpz
 * (crypto/NFLlib.cpp:504-532) and the low A bits of each coefficient packed
 * little-endian (serializeData64/32, NFLlib.cpp:412-487), as
 * PIRReplyExtraction_internal (pir/replyExtraction/PIRReplyExtraction_internal.cpp:41-106)
 * decrypts each level.  A <= 128.  Asynchronous on the context stream. */
int xpg_decrypt_dev(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *sk_dev, const uint64_t *ct_dev,
                    uint64_t n_ct, uint8_t *plain_dev);
/* query for element `index` over dims[0..d-1], noise stream `noise_seed`, on the GPU */
int xpg_query_gen_dev(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *sk_dev, uint32_t d,
                      const uint32_t *dims, uint64_t index, uint64_t noise_seed,
                      uint64_t *query_dev);
/* bytes [offset, offset+nbytes) of the seeded byte stream `seed`, written on the GPU */
int xpg_random_bytes_dev(xpg_ctx *ctx, uint64_t seed, uint64_t offset, uint8_t *dst_dev,
                         uint64_t nbytes);

#ifdef __cplusplus
}
#endif
#endif
