/*
 * admatch.h -- C ABI of the MI355X scan engine.
 *
 * Every entry point is plain C: opaque handle, plain pointers and sizes,
 * int status codes (0 = success, negative = error; ita_last_error() gives the
 * message of the calling thread's last failure).  The C++ surface of the
 * reference, include/ITA_regex.h, is implemented on top of these calls
 * (see ITA_regex.h in this directory); FFI bindings (ctypes, cgo, JNI) bind
 * these symbols directly.
 *
 * Handles are independent: each owns its compiled rule set, device tables,
 * HIP stream and buffers, so calls on distinct handles may run on distinct
 * threads.  Calls on one handle must not overlap.
 */
#ifndef ADMATCH_H
#define ADMATCH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define ITA_OK 0
#define ITA_ERR_ARG (-1)
#define ITA_ERR_LOAD (-2)
#define ITA_ERR_HIP (-3)
#define ITA_ERR_STATE (-4)
#define ITA_ERR_CAPACITY (-5)
#define ITA_ERR_TOO_LARGE (-6)
#define ITA_ERR_BUDGET (-7)
/* ita_collect: the scan's accepted states overflowed the device pool; the
 * pool has been grown to fit them, so the same scan repeated succeeds. */
#define ITA_ERR_POOL (-8)

/* Flag bits; values 1..32 are those of the reference, include/ITA_regex.h:10-24. */
#define ITA_INFA_KERNEL 1
#define ITA_AS_KERNEL (1 << 1)
#define ITA_TKO_KERNEL (1 << 2)
#define ITA_PROFILER_MODE (1 << 3)
#define ITA_SHOW_RESULTS (1 << 4)
#define ITA_AD_MATCHER (1 << 5)
/* New bits (no counterpart in the reference). */
#define ITA_LEARNED_SELECTOR (1 << 6) /* with AD_MATCHER: MLP step selector */
#define ITA_HEURISTIC_SELECTOR (1 << 7) /* with AD_MATCHER: AdMatch rule + iNFA cost term */

/* Step-selector modes of the engine kernel (ita_scan_device's mode). */
#define ITA_MODE_INFA 0
#define ITA_MODE_TKO 1
#define ITA_MODE_AS 2
#define ITA_MODE_ADMATCH 3
#define ITA_MODE_HEURISTIC 4
#define ITA_MODE_LEARNED 5
#define ITA_MODE_LOG 6
/* ita_scan_device's kind: the plain single-kernel iNFA path, or the engine. */
#define ITA_KIND_PLAIN 0
#define ITA_KIND_ENGINE 1

typedef struct ita_handle ita_handle;

/* Parse a .nfa file and build the per-method tables on the host.
 * Replaces the constructors ita_scratch::ita_scratch and
 * ad_scratch::ad_scratch (reference src/Interface.cu:20-40, graph build
 * src/transition_graph.cpp:82-365).  flags: exactly one of INFA/AS/TKO
 * kernel bits, or AD_MATCHER (reference FLAG_VERIFICATION, Interface.cu:12-18);
 * otherwise ITA_ERR_ARG.  A load failure returns ITA_ERR_LOAD (the reference
 * calls exit(-1), Interface.cu:28-31). */
int ita_create(uint16_t flags, const char* nfa_path, ita_handle** out);

/* Upload the tables to the GPU (reference allocScratch, Interface.cu:42-116).
 * It also runs a one-wave check that the engine kernel finds its arguments
 * where it reads them; ITA_ERR_HIP if not (or on any HIP error). */
int ita_alloc(ita_handle* h);

/* Release device memory (reference freeScratch, Interface.cu:124-142). */
int ita_free(ita_handle* h);

/* Destroy the handle (host tables too; the reference leaks them). */
int ita_destroy(ita_handle* h);

/* Scan one NUL-terminated string, length strlen(text) (reference Scan,
 * Interface.cu:150-172).  Writes up to cap sorted unique rule ids to out_ids
 * and the full count to *n_out; ITA_ERR_CAPACITY if it exceeds cap. */
int ita_scan(ita_handle* h, const char* text, int32_t* out_ids, int32_t cap, int32_t* n_out);

/* Scan n strings from host memory (reference BatchedScan, Interface.cu:174-198,
 * driver run_iNFA/run_TKO/run_AS/run_AD, e.g. src/iNFA_kernel.cu:175-343).
 * Rule ids of string i: out_ids[out_offsets[i] .. out_offsets[i+1]), sorted
 * unique.  out_offsets holds n+1 entries; out_ids holds cap entries.  If
 * the total exceeds cap, out_offsets is still filled and ITA_ERR_CAPACITY is
 * returned (retry with cap = out_offsets[n]). */
int ita_batched_scan(ita_handle* h, const uint8_t* const* text, const int32_t* len, int32_t n,
                     int32_t* out_offsets, int32_t* out_ids, int64_t cap);

/* Scan n strings already in device memory: d_text holds the concatenated
 * bytes and d_offsets (int64, n+1) the string boundaries.  kind: ITA_KIND_*;
 * mode: ITA_MODE_* (engine only).  stream: a hipStream_t, or NULL for the
 * handle's own stream.  use_graph != 0 captures the launch sequence in a
 * hipGraph on first use and replays it for the same (buffers, n, kind,
 * mode).  Results stay on the device until ita_collect.  No reference
 * counterpart: the reference always copies inputs per call. */
int ita_scan_device(ita_handle* h, const uint8_t* d_text, const int64_t* d_offsets, int32_t n, int32_t kind,
                    int32_t mode, void* stream, int32_t use_graph);

/* Allocation-free variant of ita_scan_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc, hipFree or other allocating routine
 * and creates no stream: it enqueues the scan on `stream` (NULL: the null
 * stream) and returns without waiting.  The kernels' counters and
 * longest-first order live in a fixed device array of the library, about
 * 4 MiB per GPU, held until the process ends; n may not exceed 2^20
 * (ITA_ERR_TOO_LARGE).  Calls that overlap on one GPU share that array and
 * race: serialise them.  kind/mode as ita_scan_device; the engine runs with
 * LDS-bitmap storage whatever ita_set_variant's store says; the handle must
 * be allocated (ita_alloc).  Outputs, all device memory of the caller: the
 * accepted states of string i, ascending, in
 * d_states[d_starts[i] .. d_starts[i] + d_counts[i]) (d_counts[i] < 0: the
 * list did not fit in states_cap entries); d_status[0] = 1 if any list did
 * not fit, d_status[1] = accepted states of all strings.  Rule ids of a
 * state come from the accept map (ita_table 21 and 22). */
int ita_scan_device_noalloc(ita_handle* h, const uint8_t* d_text, const int64_t* d_offsets, int32_t n, int32_t kind,
                            int32_t mode, void* stream, int32_t* d_counts, int64_t* d_starts, uint32_t* d_states,
                            int64_t states_cap, int64_t* d_status);

/* Copy the last ita_scan_device results to the host and map accepted states
 * to rule ids (layout as ita_batched_scan). */
int ita_collect(ita_handle* h, int32_t n, int32_t* out_offsets, int32_t* out_ids, int64_t cap);

/* Streaming scans (no counterpart in the reference, which restarts every
 * scan from the initial states, src/iNFA_kernel.cu:41-44): a string arrives
 * as chunks over several calls and the scan carries its state between them.
 * ita_stream_create allocates device state for n_streams streams of handle h
 * (valid while h stays allocated); ita_stream_scan scans chunk i as the
 * continuation of stream stream_ids[i] (each stream at most once per call)
 * and returns, per chunk, the stream's accepted rule ids as if the stream
 * ended after that chunk: the chunks of a stream scanned in order give the
 * results of their concatenation.  The word-boundary decision at a chunk's
 * last byte is deferred to the next chunk.  On ITA_ERR_CAPACITY the streams
 * have advanced and out_offsets holds the counts.  ita_stream_reset starts a
 * stream anew. */
typedef struct ita_streams ita_streams;
int ita_stream_create(ita_handle* h, int32_t n_streams, ita_streams** out);
int ita_stream_reset(ita_streams* s, int32_t stream);
int ita_stream_scan(ita_streams* s, const uint8_t* const* chunks, const int32_t* len, const int32_t* stream_ids,
                    int32_t n, int32_t* out_offsets, int32_t* out_ids, int64_t cap);
int ita_stream_destroy(ita_streams* s);
/* The rule ids of the handle's last scan or stream scan (all strings,
 * concatenated; the offsets were returned by that call): copies
 * min(total, cap) and returns the total. */
int64_t ita_last_results(ita_handle* h, int32_t* out_ids, int64_t cap);

/* Kernel time (ms, hipEvents) of the last scan and steps per method
 * (iNFA, TKO, AS) of the last engine scan. */
int ita_last_stats(ita_handle* h, float* kernel_ms, uint64_t* method_steps3);

/* The kernel instance of the handle's last scan (no counterpart in the
 * reference; introspection): out2[0] = the engine kernel's mode argument,
 * the scan mode plus 0x100 for a lean kernel (rule set without word-boundary
 * transitions, with two-byte rows, not a stream: the boundary, stream and
 * missing-row code compiled out) and 0x200 for the narrow register profile,
 * or -1 for the plain iNFA kernel; out2[1] = 1 for hashed-list storage. */
int ita_last_kernel(ita_handle* h, int32_t* out2);

/* Times (ms) of the handle's last scan: [0] the launch sequence (as
 * ita_last_stats), [1] the engine kernel alone (hipEvents around it; a
 * graph-captured scan replays its prefix kernels as a graph and launches the
 * engine kernel after it, so this holds with use_graph too; 0 for the plain
 * kernel), and for the last
 * host-strings scan [2] device buffer growth, [3] host-to-device copies,
 * [4] result copies, [5] the host's accept-map lookup. */
int ita_last_times(ita_handle* h, float* out6);

/* Engine launch variant: waves (strings) per workgroup in {1,2,4}, active-list
 * capacity in {128,256,512,1024}; plain_bs in {64,...,1024} for the plain kernel;
 * store selects the AS and learned modes' state storage: 0 = two S-bit
 * bitmaps per string in LDS, 1 = active lists with LDS hash tables (LDS per
 * string independent of S) and the S-bit bitmaps in device memory, used by
 * dense steps only.  The other modes always use LDS bitmaps. */
int ita_set_variant(ita_handle* h, int32_t wpb, int32_t cap, int32_t plain_bs, int32_t store);

/* Grouping threshold of the TKO shift groups (src/transition_graph.cpp:298-357
 * fixes it at state_count/16): an offset group of a byte becomes a shift row
 * when it holds more than `threshold` transitions, the rest are scattered as
 * single edges.  -1 = the reference's state_count/16, -2 = auto (the
 * default): W/2 with W = ceil(state_count/64), since a shift row touches W
 * words at about the cost of W/2 scattered edges; >= 0 = that count.  The
 * partition changes only how TKO steps run, never results.  Call before
 * ita_alloc (ITA_ERR_STATE after it); ita_info's [5] reports the count in
 * use. */
int ita_set_tko_threshold(ita_handle* h, int32_t threshold);
int ita_get_variant(ita_handle* h, int32_t* wpb, int32_t* cap, int32_t* plain_bs, int32_t* store);

/* Learned step selector: 5,443 packed float weights (layout in
 * admatch_amd/csrc/device.hpp).  every == 0 (default): the network is
 * evaluated on the GPU once per (active-count bucket, list validity, byte)
 * into a policy table (table 16 of ita_table) that every step reads;
 * every > 0: evaluate the network in the scan every `every` steps.  weights == NULL removes the selector (the heuristic
 * is used). */
int ita_set_selector(ita_handle* h, const float* weights, int32_t n, int32_t every);

/* Replaces the selector's policy table (table 16 of ita_table: n = 512
 * words, layout there) with one the host derived from the network, such as
 * the method of least predicted cost per entry summed over logged steps
 * (admatch_amd/selector.py policy_from_log).  Needs ita_set_selector first;
 * ita_set_selector rebuilds the table from the network again.  No
 * counterpart in the reference. */
int ita_set_policy(ita_handle* h, const uint32_t* table, int32_t n);

/* Step logging for selector training (ITA_MODE_LOG): room for cap records
 * of 22 floats (16 features, cycles of iNFA, TKO and AS or -1, the
 * heuristic's method, the byte, the explicit active count), one step
 * sampled in every `stride`. */
int ita_set_log(ita_handle* h, int64_t cap_records, int32_t stride);
int64_t ita_fetch_log(ita_handle* h, float* out, int64_t cap_records);

/* Rule-set facts: [0] states, [1] u64 words per vector, [2] transitions,
 * [3] word-boundary transitions, [4] TKO groups, [5] TKO threshold,
 * [6] AS row entries, [7] engine LDS bytes per workgroup (current variant),
 * [8] plain-kernel LDS bytes, [9] CUs (after ita_alloc, else 0),
 * [10] content hash. */
int ita_info(ita_handle* h, int64_t* out11);

/* Host copies of the tables, for tests: which = 0 iNFA offsets (258 int32),
 * 1 iNFA list (2*T uint32), 2 TKO k prefix (258 int32), 3 TKO shifts
 * (int32), 4 TKO lim (u64, K*W), 5 TKO leftover offsets (258 int32),
 * 6 TKO leftover (2*n uint32), 7 AS offsets (S*256+1 uint32), 8 AS dst
 * (uint32), 9 init/persistent/accept vectors (3*W u64), 10 device AS
 * table (S*256 uint32, entry of state s on byte c at s*256 + c: row length
 * in the top 8 bits, 1 = the low 24 bits are the destination, 254 = the
 * row is s itself and the low 24 bits, 255 = long row whose length is the
 * store word at the low 24 bits, followed by the row, otherwise the row
 * starts at that store index), 11 its multi-row store (uint32),
 * 12 always-active row offsets (257 int32),
 * 13 always-active rows (uint32), 14 two-byte row offsets (65537 uint32),
 * 15 two-byte rows (uint32), 16 the learned selector's policy table
 * (512 uint32, after ita_set_selector: method of (bucket b, list valid v,
 * byte c) in bits 2*(c%16).. of word (2*b + v)*16 + c/16, bucket =
 * min(15, floor(log2(explicit count + 1)))), 17 word-boundary successors
 * per state: offsets (S+1 uint32), 18 their destinations (uint32), 19
 * word-boundary successors of the implicit sets: offsets (258 int32, row c
 * for the always-active row of byte c, row 256 for the always-active states
 * alone), 20 their destinations (uint32); 17-20 are empty unless no
 * word-boundary destination is a word-boundary source; 21 accept map offsets
 * (S+1 int32) and 22 its rule ids (int32): the rules of state s are
 * [off[s], off[s+1]).  Returns the element
 * count; copies min(count, cap) elements when out != NULL. */
int64_t ita_table(ita_handle* h, int32_t which, void* out, int64_t cap);

/* Learned-selector training on the GPU (no counterpart in the reference):
 * samples are 16 features, 3 per-method targets (log2 step cost) and 3
 * mask values (1 = the method was timed).  Weights use the packed layout
 * of ita_set_selector.  ita_selector_train runs Adam on shuffled minibatches
 * (weights_inout: initial in, trained out; loss_history[epochs], optional);
 * ita_selector_grad returns the loss gradient at the given weights;
 * ita_selector_forward evaluates the network (out[n][3]). */
int ita_selector_train(const float* features, const float* targets, const float* mask, int64_t n, int32_t epochs,
                       int32_t batch, float lr, uint64_t seed, float* weights_inout, float* loss_history);
int ita_selector_grad(const float* features, const float* targets, const float* mask, int64_t n,
                      const float* weights, float* grad_out, float* loss_out);
int ita_selector_forward(const float* features, int64_t n, const float* weights, float* out);

/* Serving loop (no counterpart in the reference): replays n requests
 * (buf[offsets[i] .. offsets[i+1]), arriving at arrival_s[i] seconds after
 * the start, non-decreasing) in real time with continuous batching:
 * `lanes` batches in flight on their own streams; a free lane takes every
 * arrived request not yet started, up to max_batch strings and
 * max_batch_bytes input bytes, while the device bytes of all in-flight
 * batches stay within state_budget_bytes (ITA_ERR_BUDGET if one request
 * alone exceeds it); requests up to split_bytes long and longer ones wait in
 * separate queues and are batched apart, long batches holding at most half
 * of the lanes (0: one queue).  Per request: latency_s[i] (completion -
 * arrival) and rule ids in out_ids[out_offsets[i] .. out_offsets[i+1]).  stats[0..5]:
 * wall seconds, batches, mean batch size, peak in-flight device bytes, peak
 * busy lanes, total input bytes. */
int ita_serve_trace(ita_handle* h, int32_t kind, int32_t mode, int32_t lanes, int32_t max_batch,
                    int64_t max_batch_bytes, int64_t state_budget_bytes, int64_t split_bytes, const uint8_t* buf,
                    const int64_t* offsets, const double* arrival_s, int32_t n, double* latency_s,
                    int32_t* out_offsets, int32_t* out_ids, int64_t cap, double* stats6);

/* Device-side serving (no counterpart in the reference): replays n requests
 * (buf[offsets[i] .. offsets[i+1]), arriving arrival_s[i] seconds after the
 * start, non-decreasing) through ONE persistent engine launch.  The host
 * publishes each request when it arrives; a wave that finishes a request
 * takes the next published one, copies its bytes from host memory and scans
 * it, and writes the accepted states and a done flag to host memory, which
 * the host polls.  Every wait in the kernel is bounded (last arrival + 10 s).
 * The launch holds the whole GPU until the last request completes; run no
 * other work on the device meanwhile.  mode: ITA_MODE_* (engine).  Per
 * request: latency_s[i] (completion seen by the host - arrival) and rule ids
 * in out_ids[out_offsets[i] .. out_offsets[i+1]).  stats6: [0] seconds from
 * the first arrival to the last completion, [1] requests scanned again after
 * the accepted-state buffer overflowed, [2] waves in the launch, [3..5] steps
 * per method (iNFA, TKO, AS). */
int ita_serve_ring(ita_handle* h, int32_t mode, const uint8_t* buf, const int64_t* offsets, const double* arrival_s,
                   int32_t n, double* latency_s, int32_t* out_offsets, int32_t* out_ids, int64_t cap, double* stats6);

/* Message of this thread's last failed call. */
const char* ita_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* ADMATCH_H */
