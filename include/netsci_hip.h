/*
 * netsci_hip.h — the C-ABI of libnsx.so, the MI355X (gfx950) engine behind
 * the netcalc / cuarray API.
 *
 * Conventions:
 *   - Every function returns 0 on success and a non-zero status otherwise;
 *     nsx_last_error() returns a thread-local message for the last failure.
 *     No C++ exceptions cross this boundary.
 *   - Pointers named d_* are device (HBM) pointers, h_* are host pointers.
 *   - `stream` is a hipStream_t passed as void*; NULL is the default stream.
 *     Entry points that take a stream are stream-ordered and return before
 *     the work finishes unless stated otherwise.
 *   - The compute entry points share per-process scratch (grow-only device
 *     workspaces, the table buffer, a pinned flag block): issue them from one
 *     host thread at a time, and do not overlap calls on different streams;
 *     work queued on one stream runs in order.
 *   - X holds one node per row. Within a row, coordinate dimension di of
 *     frame f sits at X[row*ld + di*dstride + f] — the reference's layout
 *     (netchem/src/network.cpp:65-69, mutual_information_cpu.cpp:27-45
 *     indexing `i + di*n`), so for a plain call dstride == n.
 */
#ifndef NETSCI_HIP_H
#define NETSCI_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status and device ---------------------------------------------- */

/* Message of the calling thread's last failure ("" if none). */
const char *nsx_last_error(void);

/* Number of visible HIP devices (0 without a GPU; never fails). */
int nsx_device_count(int *count);

/* Make `device` current for the calling thread and check it is gfx950. */
int nsx_init(int device);

/* Device name and CU count, for run records. name gets at most len bytes. */
int nsx_device_info(char *name, int len, int *cu_count, int *clock_khz);

/* ---- buffers: the HIP allocator behind CuArray --------------------------
 * Replaces cudaMalloc / cudaFree / cudaMemcpy in CuArray<T>::allocateDevice,
 * deallocateDevice, toDevice and toHost (cuarray/src/cuarray.cu:104-164).
 * nsx_memcpy_h2d / d2h copy synchronously through a pinned staging ring,
 * so pageable host buffers transfer at pinned speed. */
int nsx_alloc(void **d_ptr, size_t bytes);
int nsx_free(void *d_ptr);
int nsx_host_alloc(void **h_ptr, size_t bytes);   /* pinned host memory */
int nsx_host_free(void *h_ptr);
int nsx_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes);
int nsx_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes);
int nsx_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes, void *stream);
int nsx_stream_sync(void *stream);
int nsx_device_sync(void);

/* Device timing on `stream` (hipEvent pairs). */
int nsx_event_create(void **ev);
int nsx_event_destroy(void *ev);
int nsx_event_record(void *ev, void *stream);
int nsx_event_elapsed_ms(void *start, void *stop, float *ms);

/* ---- KSG mutual information over node pairs ---------------------------
 * Replaces the per-pair GPU path of the pair loops
 * netcalc/src/mutual_information.cpp:103-172 and
 * netcalc/src/generalized_correlation.cpp:104-173, i.e. one
 * mutualInformationGpu call (netcalc/src/mutual_information_gpu.cu:1007-1079)
 * per ab row, with a single persistent launch for all P pairs.
 *
 * d_mi_out[p] = KSG MI of rows (d_ab[2p], d_ab[2p+1]) over frames 0..n-1,
 * equal bit-for-bit to mutualInformationCpu
 * (netcalc/src/mutual_information_cpu.cpp:6-121), ties at the k-th
 * neighbour included. 1 <= d <= 3, 1 <= k <= 16, k < n.
 * d_counts_out (nullable) receives n_x, n_y per frame: [p][f][2] int32.
 * Coordinates are expected finite: with a NaN coordinate the marginal max
 * here drops the NaN, where the reference's std::max keeps it when it sits
 * in the first dimension. */
int nsx_ksg_mi(const float *d_X, long long ld, long long dstride, int d,
               int n, const int *d_ab, long long P, int k, float *d_mi_out,
               int *d_counts_out, void *stream);

/* Windowed (bootstrap-replicate) variant: replicate r uses frames
 * [h_win_off[r], h_win_off[r] + h_win_n[r]) of every row;
 * d_mi_out[r*P + p]. All R*P (replicate, pair) items run in one persistent
 * launch, longest windows first. No counterpart in the reference, whose
 * bootstrap resamples and calls the pair loop once per replicate
 * (docs/tutorial3_bootstrapping.rst:125); each item equals that call on the
 * window's frames. */
int nsx_ksg_mi_windows(const float *d_X, long long ld, long long dstride,
                       int d, const int *h_win_off, const int *h_win_n, int R,
                       const int *d_ab, long long P, int k, float *d_mi_out,
                       void *stream);

/* Allocation-free variant of nsx_ksg_mi (d_counts_out = NULL). No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same results as nsx_ksg_mi, computed by
 * the all-frames kernel, whose scratch (psi table, tie counter, per-wave tie
 * rows) lives in a 64 MiB static device array that each GPU holds until the
 * process ends; it calls no hipMalloc / hipFree. n <= 262144 frames (larger
 * n is rejected). The launch is capped to the waves whose tie rows fit
 * (64 MiB / (8 n) waves). Stream-ordered: it enqueues and returns without
 * synchronising, and does not update nsx_ksg_last_tie_frames. Every call on
 * a GPU shares that GPU's array, so calls that overlap on one GPU race. */
int nsx_ksg_mi_noalloc(const float *d_X, long long ld, long long dstride, int d,
                       int n, const int *d_ab, long long P, int k, float *d_mi_out,
                       void *stream);

/* The pair loop in chunks, for checkpointing callers: the reference's
 * checkpointed pair loops (netcalc/src/generalized_correlation.cpp:8-101,
 * mutual_information.cpp:8-101) save R after every checkpointFrequency
 * pairs. Same results as nsx_ksg_mi over all P pairs, computed in chunks
 * [h_ends[c-1], h_ends[c]) (h_ends strictly increasing, the last equal to P)
 * with the list path's tables built once. When chunk c is done, its MI
 * values are in d_mi_out and h_mi_out[h_ends[c-1] .. h_ends[c]), and
 * on_chunk(user, h_ends[c]) runs on the calling thread while the device
 * works on the next chunks; a non-zero return stops the call with an error.
 * Consecutive chunks of fewer than 128 pairs per CU in all are computed
 * in one device batch, and their callbacks run in order when it is done.
 * Returns when every chunk is done. */
typedef int (*nsx_chunk_fn)(void *user, long long end);
int nsx_ksg_mi_chunked(const float *d_X, long long ld, long long dstride, int d,
                       int n, const int *d_ab, long long P, int k, float *d_mi_out,
                       float *h_mi_out, const long long *h_ends, int nends,
                       nsx_chunk_fn on_chunk, void *user, void *stream);

/* Frames whose k-th / (k+1)-th joint distances tied in the last KSG launch
 * on this thread (resolved on the device in libstdc++ std::sort order). */
int nsx_ksg_last_tie_frames(long long *count);

/* Launch shape for the eager KSG calls above: queries per lane (1 or 2),
 * LDS frame tile (64, 128 or 256), waves per workgroup (1..8), workgroups per
 * CU (0 = as many as fit), and symmetric = 1 for the pass-2 variant that
 * evaluates each unordered frame pair once and credits both frames (ballot +
 * popcount). qpl = 0 clears the override (tuned or default shapes apply
 * again). Results do not depend on the shape. No reference counterpart (its
 * launch shape is fixed, mutual_information_gpu.cu:1055-1062). */
int nsx_ksg_set_config(int qpl, int chunk, int wpb, int blocks_per_cu, int symmetric);

/* Algorithm of nsx_ksg_mi and nsx_ksg_mi_windows (each window on its own
 * tables, one after another): method 0 = the sorted-list path
 * whenever it applies (at most 8192 frames whose two rows fit in LDS, finite
 * coordinates, room for the tables of 4*n*n bytes of at least two nodes:
 * when all used nodes do not fit, they are built block by block) and the
 * all-frames kernel otherwise; 1 = always the all-frames kernel; 2 = the list path or
 * an error. walk_waves (1..16) and walk_batch (1..64) shape the list path's
 * walk kernel (0 = default). Results are identical for every choice. No
 * reference counterpart (the reference always compares all frame pairs,
 * mutual_information_gpu.cu:62-687). */
int nsx_ksg_set_method(int method, int walk_waves, int walk_batch);

/* Which algorithm the last nsx_ksg_mi call on this thread ran (1 all-frames,
 * 2 sorted lists) and, for the list path, the list entries walked (both
 * lists, summed over all query frames). */
int nsx_ksg_last_method(int *method, long long *list_entries);

/* Device time (ms, hipEvents on the call's stream) of the last list-path
 * nsx_ksg_mi call's phases: phase_ms[0] node map and table build, [1] walk,
 * [2] tie resolution, [3] frame-order psi sums. A diagnostic for roofline
 * accounting; no reference counterpart. */
int nsx_ksg_last_phase_ms(double *phase_ms);

/* Host wall time (ms) of the last nsx_ksg_mi / nsx_ksg_mi_windows call on
 * this thread, from entry to return (it ends with a stream synchronisation),
 * i.e. the device phases above plus the host's own work. No reference
 * counterpart. */
int nsx_ksg_last_call_ms(double *wall_ms);

/* The list path's table layout in the last call on this thread: blocks = 1
 * when every used node's tables fit at once, else the node blocks of the
 * blocked schedule (one block resident while the others stream through the
 * second half of the tables); slots = table slots held; node_builds = node
 * tables built, summed over windows. Zeros when the call did not take the
 * list path. No reference counterpart. */
int nsx_ksg_last_list_info(int *blocks, long long *slots, long long *node_builds);

/* Cap (bytes) on the device memory the list path's tables may use; 0 = the
 * default, 85% of free device memory less 1 GiB. The tables are kept between
 * calls and grow only when a call needs more; a cap below what is held
 * applies from the next call. No reference counterpart. */
int nsx_ksg_set_table_budget(long long bytes);

/* Entries kept per sorted list of the list path's tables (rounded up to 8,
 * at least 8); 0 = the default, all of a row up to 512 frames and max(512,
 * n / 8) beyond. Walks that reach the end of a cut list without meeting
 * their stopping rule are resolved exactly by the tie kernels, so any cap
 * gives the same results; a smaller one saves memory. Applies from the next
 * call. No reference counterpart. */
int nsx_ksg_set_list_cap(long long entries);

/* Frees the library's device workspaces and tables (synchronises the
 * device); the next call allocates what it needs again. No reference
 * counterpart. */
int nsx_release_workspace(void);

/* ---- batched plans: one hipGraph launch per replay ----------------------
 * A plan fixes X, the windows, the pairs and the output of
 * nsx_ksg_mi_windows, owns every buffer its graph uses, and captures the
 * whole pass into a hipGraph, so a replay is one hipGraphLaunch with no
 * host synchronisation: on the sorted-list path (nsx_ksg_plan_method = 2)
 * every window's table build, walk, tie resolution and psi sums; on the
 * all-frames kernel [tie-counter reset, persistent KSG launch]. autotune != 0 times candidate launch shapes
 * (queries per lane x LDS frame tile x waves per workgroup) on a sample of
 * the items first and keeps the fastest (cached per (d, k, window-length
 * bucket); persisted to the file named by NSX_TUNE_CACHE when set).
 * No reference counterpart: the reference runs one pair loop per bootstrap
 * replicate (docs/tutorial3_bootstrapping.rst:125). */
typedef struct nsx_ksg_plan_s *nsx_ksg_plan;
int nsx_ksg_plan_create(nsx_ksg_plan *plan, const float *d_X, long long ld,
                        long long dstride, int d, const int *h_win_off,
                        const int *h_win_n, int R, const int *d_ab, long long P,
                        int k, float *d_mi_out, int autotune);
int nsx_ksg_plan_launch(nsx_ksg_plan plan, void *stream);
/* Tie frames of the last replay (synchronises the device). */
int nsx_ksg_plan_tie_frames(nsx_ksg_plan plan, long long *count);
int nsx_ksg_plan_config(nsx_ksg_plan plan, int *qpl, int *chunk, int *wpb,
                        long long *blocks, double *tune_ms, int *symmetric);
int nsx_ksg_plan_destroy(nsx_ksg_plan plan);
/* Algorithm a plan's graph runs: 2 = the sorted-list path (chosen, like
 * nsx_ksg_set_method's automatic choice, when the windows have at most 8192
 * frames whose two rows fit LDS and the coordinates are finite at creation;
 * the graph then rebuilds every window's tables on each replay and carries
 * the all-frames launch as a fallback that runs only if a coordinate has
 * become non-finite), 1 = the all-frames kernel. */
int nsx_ksg_plan_method(nsx_ksg_plan plan, int *method);

/* GC transform of generalized_correlation_{cpu,gpu}.cpp:8-33 on host data:
 * h_gc[p] = MI <= 0 ? 0 : (float)sqrt(1 - (float)exp(-(2/(float)d)*MI)),
 * evaluated in double like the reference. */
int nsx_gc_from_mi_host(const float *h_mi, long long P, int d, float *h_gc);

/* psi table of psi.cpp:6-19: h_psi[0..n]. */
int nsx_psi_host(float *h_psi, int n);

/* ---- learned pair screener ---------------------------------------------
 * No reference counterpart: the reference computes KSG for every pair. These
 * entry points let all-pairs GC run the exact KSG path only on pairs a small
 * MLP predicts may exceed a GC threshold, plus an audit sample; the values
 * reported for computed pairs are always exact KSG.
 *
 * nsx_screen_features: 16 bf16 features per pair (see screener.hip) from the
 * trajectory's covariance (d = 3, X in the layout above). d_feat: P x 16
 * uint16 (bf16 bits). */
int nsx_screen_features(const float *d_X, long long ld, long long dstride, int N,
                        int n, const int *d_ab, long long P,
                        unsigned short *d_feat, void *stream);
/* MLP 16 -> 64 -> 64 -> 1 (ReLU) with FP32 master parameters in one array of
 * nsx_mlp_param_count() floats: W1[16][64], b1[64], W2[64][64], b2[64],
 * w3[64], b3. Products run on bf16 MFMA with FP32 accumulation. */
int nsx_mlp_param_count(int *count);
int nsx_mlp_forward(const float *d_params, const unsigned short *d_feat,
                    long long P, float *d_out, void *stream);
/* Gradient of mean_i w_i (out_i - y_i)^2 over a batch of B samples (summed
 * deterministically) and the loss. */
int nsx_mlp_grad(const float *d_params, const unsigned short *d_feat,
                 const float *d_y, const float *d_w, long long B, float *d_grad,
                 float *d_loss, void *stream);
/* Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the FP32 master parameters. */
int nsx_mlp_adam(float *d_params, const float *d_grad, float *d_m, float *d_v,
                 int step, float lr, void *stream);
/* Pairs with prediction >= threshold, plus a hashed audit sample of about
 * audit_rate of the others, compacted into d_ab_out (pairs) and d_idx_out
 * (their indices); d_flag[p] = 1 candidate, 2 audit, 0 skipped. Returns the
 * selected and audit counts on the host (synchronises the stream). */
int nsx_screen_select(const float *d_pred, long long P, float threshold,
                      float audit_rate, unsigned seed, const int *d_ab,
                      int *d_ab_out, long long *d_idx_out, unsigned char *d_flag,
                      long long *h_n_selected, long long *h_n_audit, void *stream);

/* ---- host (CPU platform) implementations -------------------------------
 * The CPU side of the `platform` flag (platform.h:9-10). Same results as the
 * GPU path; threads over pairs (threads <= 0: OMP_NUM_THREADS when set,
 * else every hardware thread; the same for the other CPU-platform calls). */
int nsx_cpu_ksg_mi(const float *h_X, long long ld, long long dstride, int d,
                   int n, const int *h_ab, long long P, int k,
                   float *h_mi_out, int threads);

/* libstdc++-order selection used for tie resolution: h_idx[0..m) receives
 * the first m entries of std::sort(iota(n), by h_key with <), as
 * mutual_information_cpu.cpp:48-53 orders a frame's neighbours.
 * h_heap_ranges (nullable) gets the number of sub-ranges that took
 * introsort's heap-sort fallback. */
int nsx_select_prefix_host(const float *h_key, int n, int m, int *h_idx,
                           int *h_heap_ranges);

/* ---- Hedetniemi all-pairs shortest paths ------------------------------
 * Replaces hedetniemiAllShortestPathsGpu (netcalc/src/hedetniemi_gpu.cu:148-235,
 * kernels :47-107). H_{p+1}[i][j] = min_k fl(A[j][k] + H_p[k][i]) for
 * p = 0..L-1 from H_0 = A, double-buffered (Jacobi), d_Hi[p] = H_p
 * (L*N*N floats, nullable when d_paths is null), then the backtrack of
 * :70-107 into d_paths[i][j][0..L-1] (N*L*N ints). Diagonal: paths[i][i][*] = i.
 * With d_paths, N <= 92680. Both Hedetniemi entry points enqueue A's
 * neighbour lists on `stream`, then wait on the host until that work (and
 * whatever `stream` already held) has produced the sparse/dense flag
 * (polled from pinned host memory), then enqueue the L steps (and the
 * backtrack) and return without further waiting. */
int nsx_hedetniemi(const float *d_A, int N, int L, float tol, float *d_H,
                   float *d_Hi, int *d_paths, void *stream);

/* hedetniemiAllShortestPathLengthsGpu (hedetniemi_gpu.cu:109-146): H only.
 * d_H may be d_A itself (the lengths overwrite A). */
int nsx_hedetniemi_lengths(const float *d_A, int N, int L, float *d_H,
                           void *stream);

/* correlationToAdjacencyGpu (hedetniemi_gpu.cu:237-261, kernel :8-26). */
int nsx_corr_to_adj(const float *d_C, int N, float *d_A, void *stream);

/* CPU platform versions of the three above (same results). */
int nsx_cpu_hedetniemi(const float *h_A, int N, int L, float tol, float *h_H,
                       int *h_paths, int threads);
int nsx_cpu_corr_to_adj(const float *h_C, int N, float *h_A);

/* ---- row statistics (netcalc/include/math_utils.h:24-101) ---------------
 * meanGpu / standardDeviationGpu (math_utils_gpu.cu:158-206, kernels
 * :23-156): per row of an m x n float32 matrix, the mean u and the
 * population standard deviation sigma, summed in the reference's order
 * (1024 strided partial sums folded as a tree), so the values are the
 * reference kernel's bits (the square is fl(d * d), where the reference
 * calls powf(d, 2)). n >= 1. */
int nsx_row_mean(const float *d_a, long long m, long long n, float *d_u, void *stream);
int nsx_row_std(const float *d_a, long long m, long long n, float *d_u, float *d_sigma,
                void *stream);
/* CPU platform: the same sums on the host (h_sigma nullable). The
 * reference's CPU branch computes nothing (math_utils.cpp:6-35). */
int nsx_cpu_row_stats(const float *h_a, long long m, long long n, float *h_u, float *h_sigma);

#ifdef __cplusplus
}
#endif

#endif /* NETSCI_HIP_H */
