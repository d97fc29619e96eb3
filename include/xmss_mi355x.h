/*
 * xmss_mi355x.h -- additions of libxmss_mi355x beyond the reference API.
 *
 * None of these has a counterpart in cuxmss/; they are additive and leave
 * the drop-in functions of xmss.h / xmss_core.h / params.h unchanged.
 */
#ifndef XMSS_MI355X_H
#define XMSS_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Randomness for keypairs: a reproducible SplitMix64 byte stream from `seed`
 * (each 64-bit output as 8 little-endian bytes), or /dev/urandom (default).
 * The reference links a different randombytes() object for the same purpose. */
void xmss_mi355x_rng_seed(uint64_t seed);
void xmss_mi355x_rng_urandom(void);

/* 1 when a HIP device is usable, 0 otherwise (then every hashing call
 * returns -3). */
int xmss_mi355x_device_ok(void);

/* Error behaviour of every hashing entry point (the reference functions of
 * xmss.h / xmss_core.h included): a failed HIP call or a device allocation the
 * GPU (or the limit below) cannot hold returns -3 and leaves the process
 * running.  A one-call signature that fails that way leaves its sk as it was
 * (no signature left the call); a batch job's inputs that reached the device
 * return -3, and their keys' indices stay advanced.  Messages longer than
 * 0xFFFFF000 bytes are refused with -1 before anything is written (sign: sk
 * unchanged; verify: *mlen = 0); the reference would hash them.
 *
 * Device memory: xmss_mi355x_device_bytes() is what the library holds now
 * (workspaces, key pools, resident serving state); set_device_limit caps it
 * (0 = no cap, the default) and returns the previous cap.  An allocation
 * beyond the cap fails as above. */
size_t xmss_mi355x_set_device_limit(size_t bytes);
size_t xmss_mi355x_device_bytes(void);

/* Verification that allocates no device memory.  No counterpart in the
 * reference: for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Same
 * arguments, return values and m / *mlen behaviour as xmss_sign_open
 * (mt = 0) or xmssmt_sign_open (mt = 1), computed by the same kernels, but
 * every device buffer of the call -- staged inputs, key context, pre-images,
 * digests, roots -- lives in one fixed 16 MiB device array per GPU that the
 * library holds until the process ends; the inputs are copied from the
 * caller's buffers as they are.  A request needing more than the array holds
 * (messages of about 15 MB and more) returns -4 (with *mlen = 0 and a message
 * on stderr).  Calls are serialized by the library; the array keeps the last
 * call's contents, and the blob region is cleared before each use. */
int xmss_mi355x_sign_open_static(int mt, unsigned char *m, unsigned long long *mlen,
                                 const unsigned char *sm, unsigned long long smlen,
                                 const unsigned char *pk);

/* Per-kernel timing with HIP events on the library's stream. */
void xmss_mi355x_profile(int on);
void xmss_mi355x_reset_stats(void);
/* Writes "name count total_ms\n" lines; returns the length needed. */
size_t xmss_mi355x_kernel_stats(char *buf, size_t len);

/* Measures the int32 VALU issue rate of the device: `blocks` workgroups of
 * 256 lanes each run iters x 96 independent 32-bit integer instructions
 * (the SHA-256 round mix).  Writes lane-ops/s and the kernel time; returns
 * 0, or -3 without a GPU.  Used as the measured peak of the roofline. */
int xmss_mi355x_valu_peak(unsigned blocks, unsigned iters, double *lane_ops_per_s, double *ms);
/* The same for one instruction class: kind 0 is the mix above, 1 v_alignbit_b32,
 * 2 v_add3_u32, 3 v_bitop3_b32, 4 v_add_u32 / v_xor_b32, 5 v_perm_b32 / v_bfi_b32
 * (96 instructions per lane and iteration each).  Returns -1 for an unknown kind. */
int xmss_mi355x_valu_probe(int kind, unsigned blocks, unsigned iters, double *lane_ops_per_s, double *ms);

/* ---- batches of independent keys, one parameter set per batch ----------
 * Plain-core signing of `count` OID-prefixed secret keys (each updated in
 * place exactly as xmss[mt]_sign would) and verification of `count`
 * signed messages.  The three-step form stages inputs in HBM (prepare),
 * computes with no host traffic (run) and copies results out (collect);
 * rets[i] carries what xmss[mt]_sign / _sign_open would have returned.
 * The pointer arrays are copied by prepare; the buffers they point to (keys,
 * messages, signatures) must stay valid until collect returns. */
typedef struct xmss_mi355x_job xmss_mi355x_job;

xmss_mi355x_job *xmss_mi355x_sign_prepare(int mt, unsigned char *const *sks,
                                          const unsigned char *const *ms,
                                          const unsigned long long *mlens,
                                          size_t count);
xmss_mi355x_job *xmss_mi355x_verify_prepare(int mt, const unsigned char *const *sms,
                                            const unsigned long long *smlens,
                                            const unsigned char *const *pks,
                                            size_t count);
/* Mixed parameter sets in one job: input i belongs to XMSS (mts[i] = 0) or
 * XMSS^MT (mts[i] = 1) and to the parameter set its OID names; each set gets
 * its own batch on its own engine (HIP streams), and job_run queues all of
 * them before waiting, so the sets share the GPU.  collect and free as above. */
xmss_mi355x_job *xmss_mi355x_mixed_sign_prepare(const int *mts, unsigned char *const *sks,
                                                const unsigned char *const *ms,
                                                const unsigned long long *mlens, size_t count);
xmss_mi355x_job *xmss_mi355x_mixed_verify_prepare(const int *mts, const unsigned char *const *sms,
                                                  const unsigned long long *smlens,
                                                  const unsigned char *const *pks, size_t count);
/* New inputs for a verification job, of its shape: the same count, each
 * input's public key of the same parameter set (OID) as before and each
 * signed message again at least sig_bytes long (or again shorter).  The job
 * keeps its device buffers and, once run twice, the HIP graph its launches
 * were captured into, so a stream of same-shape batches replays one graph
 * per batch.  Returns 0, or -1 (job unchanged) when the shape differs. */
int xmss_mi355x_verify_refill(xmss_mi355x_job *job, const unsigned char *const *sms,
                              const unsigned long long *smlens, const unsigned char *const *pks);
/* Runs a job on the GPU (no host traffic).  A verification job's launches
 * are captured into a HIP graph on its second run and replayed afterwards.
 * job_run = job_launch + job_wait. */
int xmss_mi355x_job_run(xmss_mi355x_job *job);
/* Autotuner of the verification-batch schedule.  How a large verification
 * batch is laid onto the GPU (how its requests split into groups by message
 * length, which launches finish the L-trees and auth paths, wave priorities,
 * where the address-only work runs) is one of a few candidate plans; the
 * best depends on the parameter set, the batch size and whether the batch
 * shares the GPU with other sets.  A job of a shape not seen before times
 * every candidate on its second run -- each captured into a HIP graph and
 * replayed on the job's own inputs, every replay a whole verification -- and
 * keeps the fastest for itself and every later job of that shape (and, when
 * XMSS_MI355X_TUNE_FILE names a file, for later processes; XMSS_MI355X_TUNE=0
 * keeps the default plan).  No counterpart in the reference, whose launch
 * shapes are compile-time constants (cuxmss/gpu/gpu_verify.cu:393-502).
 * set_verify_plan forces one plan for jobs run from now on (plan >= 0) or
 * returns to tuning (-1) and returns the previous setting; verify_plans
 * writes up to max candidate plans and returns their number; tuned_plans
 * writes one line per settled shape and returns the bytes needed. */
int xmss_mi355x_set_verify_plan(int plan);
size_t xmss_mi355x_verify_plans(unsigned *plans, size_t max);
size_t xmss_mi355x_tuned_plans(char *buf, size_t len);
/* job_launch queues a job's device work and returns at once; job_wait
 * blocks until it is done.  Refill, collect and free of a launched job wait
 * for it by themselves.  A verification job's input copy (prepare, refill)
 * and a launched job's result copy run apart from the launched work, and
 * refill returns once the inputs are in pinned host memory, so with two jobs
 * a stream of batches overlaps host staging and the PCIe copies with the GPU
 * work:  launch(A); loop { refill(B); collect(A); launch(B); swap(A, B); }
 * and with three jobs in rotation the GPU also has the next batch queued
 * while the host collects.  (A signing job's input copy keeps the compute
 * stream.)  Each returns 0, -1 for a null job, or -3 when a device call
 * failed: the job is then failed, and collect reports -3 for every input that
 * reached the device (refill, too, returns -3 for a failed job). */
int xmss_mi355x_job_launch(xmss_mi355x_job *job);
int xmss_mi355x_job_wait(xmss_mi355x_job *job);
/* Collect returns 0, or -1 for a null job or one of the other kind (sign /
 * verify); free ignores a null job. */
int xmss_mi355x_sign_collect(xmss_mi355x_job *job, unsigned char *const *sms,
                             unsigned long long *smlens, int *rets);
/* ms[i] must hold smlens[i] bytes, as for xmss_sign_open. */
int xmss_mi355x_verify_collect(xmss_mi355x_job *job, unsigned char *const *ms,
                               unsigned long long *mlens, int *rets);
void xmss_mi355x_job_free(xmss_mi355x_job *job);

/* Plain-core keypairs of `count` independent keys of one parameter set from
 * caller-supplied seeds (3n bytes each: SK_SEED || SK_PRF || PUB_SEED), as
 * xmssmt_core_seed_keypair does per key, with the OID prefix of
 * xmss[mt]_keypair.  All top subtrees are built in one GPU pipeline. */
int xmss_batch_keypair(int mt, unsigned char *const *pks, unsigned char *const *sks,
                       const unsigned char *seeds, uint32_t oid, size_t count);

/* ---- one key, many messages (plain core: libxmss_mi355x.so only) --------
 * Signs ms[0..count) with consecutive indices of one OID-prefixed secret key:
 * the signatures, return values (rets[i]) and final sk are exactly those of
 * `count` successive xmss[mt]_sign calls, but every subtree the signatures
 * share is built once (one tree pass for up to 2^h signatures of an XMSS
 * key).  The counterpart of the reference's GPU one-key mode
 * (gpu/gpu_sign_origin.h:23-26, gpu_xmss[mt]_opk_sign_origin), with the
 * messages given separately.  Returns 0, -1 for an unknown OID, or -3
 * without a GPU.  xmss_batch_sign_open verifies such signatures in one
 * batch (pks[i] may all be the same key). */
int xmss_mi355x_sign_many(int mt, unsigned char *sk, unsigned char *const *sms,
                          unsigned long long *smlens, const unsigned char *const *ms,
                          const unsigned long long *mlens, size_t count, int *rets);

/* One-call forms of the above. */
int xmss_batch_sign(int mt, unsigned char *const *sks, unsigned char *const *sms,
                    unsigned long long *smlens, const unsigned char *const *ms,
                    const unsigned long long *mlens, int *rets, size_t count);
int xmss_batch_sign_open(int mt, unsigned char *const *ms, unsigned long long *mlens,
                         const unsigned char *const *sms,
                         const unsigned long long *smlens,
                         const unsigned char *const *pks, int *rets, size_t count);

/* ---- serving loop (continuous batching of BDS-core requests) -------------
 * Keys are added with their BDS-format secret key (as xmss[mt]_keypair of
 * libxmss_mi355x_fast writes it) and public key; the server owns the key from
 * then on.  Signatures of one key run in submission order and update its sk
 * exactly as xmss[mt]_sign would; requests of different keys share GPU
 * launches.  Each key's state (key context and serialized sk) stays resident
 * in HBM and is advanced there; the server keeps a byte-identical host copy.
 * The server's device memory is exactly hbm_budget_bytes: one allocation made
 * at creation, holding the keys' contexts and sks and every launch's buffers.
 * Resident keys count against it, and every request reserves the device bytes
 * its launches take against what is left and waits (first in, first out)
 * while the budget is used up; a request larger than the budget left by the
 * keys returns -3.  Results come back with their latency (submission to
 * completion). */
typedef struct xmss_mi355x_server xmss_mi355x_server;

/* Returns NULL without a GPU or when the budget cannot be allocated. */
xmss_mi355x_server *xmss_mi355x_server_create(size_t hbm_budget_bytes, int workers);
/* Returns the key id, or -1 for an unknown OID, when the budget left by the
 * keys already added cannot hold the key's resident state, or when that HBM
 * cannot be allocated. */
int xmss_mi355x_server_add_key(xmss_mi355x_server *s, int mt, const unsigned char *sk,
                               const unsigned char *pk);
/* Return request ids (0 for an unknown key). */
uint64_t xmss_mi355x_server_sign(xmss_mi355x_server *s, int key, const unsigned char *m,
                                 unsigned long long mlen);
uint64_t xmss_mi355x_server_verify(xmss_mi355x_server *s, int key, const unsigned char *sm,
                                   unsigned long long smlen);
/* Waits for a request; out receives sm (sign) or m (verify).  Returns what
 * xmss[mt]_sign / _sign_open would have returned. */
int xmss_mi355x_server_result(xmss_mi355x_server *s, uint64_t id, unsigned char *out,
                              unsigned long long *outlen, double *latency_ms);
/* Current secret key of a key id (waits for its running signature). */
int xmss_mi355x_server_get_sk(xmss_mi355x_server *s, int key, unsigned char *sk);
/* peak_bytes: the most device memory the server's arena has had in use
 * (resident keys and launch buffers together), never more than the budget. */
void xmss_mi355x_server_stats(xmss_mi355x_server *s, uint64_t *launches, uint64_t *calls,
                              size_t *peak_bytes, uint64_t *admission_waits);
/* Device memory of the server: its budget (the one arena it allocated), the
 * bytes its keys hold, the bytes in use now and the most in use so far. */
void xmss_mi355x_server_memory(xmss_mi355x_server *s, size_t *budget, size_t *resident,
                               size_t *in_use, size_t *peak);
void xmss_mi355x_server_destroy(xmss_mi355x_server *s);

#ifdef __cplusplus
}
#endif

#endif
