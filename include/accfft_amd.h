/*
 * accfft_amd.h -- C ABI of the MI355X (gfx950) implementation of AccFFT's
 * single-GPU hot path.
 *
 * Plain C: opaque plan handle, explicit status codes, int64 sizes, device
 * pointers as void*.  The drop-in C++ API (accfft_gpu.h, accfft_gpuf.h,
 * accfft_operators_gpu.h, accfft_common.h in this directory) is a set of
 * inline wrappers over these entry points.
 *
 * Each declaration names the reference interface it replaces (paths are
 * relative to the reference tree).  "fp" selects the precision:
 * ACCFFT_AMD_F64 replaces the double API (accfft_gpu.h), ACCFFT_AMD_F32 the
 * float API (accfft_gpuf.h).
 *
 * Stream semantics: every call is ordered on the plan's stream (default:
 * the null stream) and returns without waiting unless the plan was created
 * with ACCFFT_AMD_BLOCKING or a timer pointer is passed (then the call
 * synchronises to read its events, as the reference does at
 * src/accfft_gpu.cpp:478-487).
 */
#ifndef ACCFFT_AMD_H
#define ACCFFT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct accfft_amd_plan accfft_amd_plan;

/* Status codes (the reference returns NULL / -1 / prints; see
 * src/accfft_gpu.cpp:186-190, :641-646 and src/accfft_common.cpp:179-184). */
enum {
  ACCFFT_AMD_OK = 0,
  ACCFFT_AMD_ERR_INVALID = 1,     /* bad argument (size <= 0, bad axis, ...) */
  ACCFFT_AMD_ERR_OVERFLOW = 2,    /* a size does not fit the target type */
  ACCFFT_AMD_ERR_UNSUPPORTED = 3, /* length, layout or mode this call does not support */
  ACCFFT_AMD_ERR_NOT_BAKED = 4,   /* plan of the wrong kind for this call */
  ACCFFT_AMD_ERR_HIP = 5,         /* a HIP runtime call failed */
  ACCFFT_AMD_ERR_ALLOC = 6,       /* device allocation failed */
  ACCFFT_AMD_ERR_NULL = 7,        /* NULL plan or pointer */
  ACCFFT_AMD_ERR_SCRATCH = 8      /* workspace above ACCFFT_AMD_NOALLOC_MAX_BYTES */
};

/* Size of the fixed device scratch of accfft_amd_inv_laplace_noalloc: one
 * half spectrum, N0*N1*(N2/2+1) complex values (160 MiB: 256^3 in FP32 or
 * FP64). */
#define ACCFFT_AMD_NOALLOC_MAX_BYTES (160u * 1024u * 1024u)

enum { ACCFFT_AMD_F64 = 0, ACCFFT_AMD_F32 = 1 };

/* Plan flags.  ACCFFT_ESTIMATE(1)/MEASURE(2)/PATIENT(4) of
 * include/accfft_common.h:40-42 (doc/flags.md): with MEASURE or PATIENT the
 * plan times the tile configurations of its strided passes on its own
 * shapes over a temporary buffer and keeps the fastest (which can change
 * results in the last bits); ESTIMATE keeps the built-in choice. */
#define ACCFFT_AMD_ESTIMATE 1u
#define ACCFFT_AMD_MEASURE 2u
#define ACCFFT_AMD_PATIENT 4u
#define ACCFFT_AMD_BLOCKING 0x100u       /* synchronise the stream after each call */
#define ACCFFT_AMD_PRESERVE_INPUT 0x200u /* c2r / inverse c2c keep their input */
/* FP32 plans only (no counterpart in the reference): radix-16 stages of
 * power-of-two lengths 256..4096 run on the matrix cores with operands split
 * into 16-bit terms and FP32 accumulation.  BF16: 2 bf16 terms (~16 bits);
 * BF16X3: 3 bf16 terms (~24 bits); FP16: 2 fp16 terms (~22 bits), each
 * butterfly scaled by a power of two.  Other lengths keep the FP32 path. */
#define ACCFFT_AMD_SPLIT_BF16 0x1000u
#define ACCFFT_AMD_SPLIT_BF16X3 0x2000u
#define ACCFFT_AMD_SPLIT_FP16 0x4000u

/* Directions (include/accfft_common.h:37-39). */
#define ACCFFT_AMD_FORWARD (-1)
#define ACCFFT_AMD_BACKWARD (+1)

const char* accfft_amd_status_string(int status);
int accfft_amd_version(void);

/* ---------------------------------------------------------------- sizes */
/* accfft_local_size_dft_r2c_gpu(f) (src/accfft_gpu.cpp:62-65,
 * src/accfft_gpuf.cpp:59-62) -> accfft_local_size_dft_r2c_t
 * (src/accfft_common.cpp:98-192) on a p0 x p1 grid at coords (c0, c1).
 * alloc_max is returned in bytes as int64 (the reference returns int). */
int accfft_amd_local_size_dft_r2c(const int n[3], int isize[3], int istart[3],
                                  int osize[3], int ostart[3], int fp, int p0,
                                  int p1, int c0, int c1, int64_t* alloc_max);
/* accfft_local_size_dft_c2c_gpu(f) (src/accfft_gpu.cpp:689-692) ->
 * accfft_local_size_dft_c2c_t (src/accfft_common.cpp:203-256). */
int accfft_amd_local_size_dft_c2c(const int n[3], int isize[3], int istart[3],
                                  int osize[3], int ostart[3], int fp, int p0,
                                  int p1, int c0, int c1, int64_t* alloc_max);

/* ---------------------------------------------------------------- plans */
/* accfft_plan_dft_3d_r2c_gpu(f) (src/accfft_gpu.cpp:77-442;
 * src/accfft_gpuf.cpp:74-441).  In place iff data_out == data (rows padded
 * to 2*(n[2]/2+1) reals).  1x1 process grid. */
int accfft_amd_plan_dft_3d_r2c(const int n[3], void* data, void* data_out, int fp,
                               unsigned flags, accfft_amd_plan** plan);
/* accfft_plan_dft_3d_c2c_gpu(f) (src/accfft_gpu.cpp:704-931). */
int accfft_amd_plan_dft_3d_c2c(const int n[3], void* data, void* data_out, int fp,
                               unsigned flags, accfft_amd_plan** plan);
/* One rank's share of accfft_plan_dft_3d_r2c_gpu(f) on a nprocs x 1 grid
 * (src/accfft_gpu.cpp:77-442 with c_dims = {nprocs, 1}): n is the global
 * size; the input is the x-slab [n0/nprocs][n1][n2] real, the output the
 * y-slab [n0][n1/nprocs][n2/2+1] complex, sizes as
 * accfft_local_size_dft_r2c_gpu (src/accfft_common.cpp:98-192) at coords 0.
 * execute_r2c runs z r2c, y c2c whose store is the pack of T_plan_2
 * (local_transpose_col, src/transpose_gpu.cpp:3889-3891), the exchange
 * (src/transpose_gpu.cpp:3940-3955) through the plan's exchange hook, and
 * the x c2c on the received blocks; execute_c2r runs the inverse (x c2c in
 * place on its input, exchange, y c2c whose load is the unpack
 * local_transpose of src/transpose_gpu.cpp:1411-1413, z c2r).  With no hook
 * the exchange is the identity (the packed blocks are taken as received):
 * the per-GPU kernels of the distributed transform at their real shapes.
 * Requires n0 % nprocs == 0, n1 % nprocs == 0, n1/nprocs a power of two,
 * and data != data_out (out of place); full xyz only. */
int accfft_amd_plan_dft_3d_r2c_slab(const int n[3], int nprocs, void* data, void* data_out,
                                    int fp, unsigned flags, accfft_amd_plan** plan);
/* Exchange hook of a slab plan (the all-to-all of T_plan_2, no counterpart
 * as a separate call in the reference): moves nprocs blocks of block_bytes
 * from send to recv, block q to rank q, on the given stream (a hipStream_t);
 * direction -1 in execute_r2c, +1 in execute_c2r.  Returns a status.  A
 * RCCL all-to-all over device buffers is the intended implementation. */
typedef int (*accfft_amd_exchange_fn)(const void* send, void* recv, int64_t block_bytes,
                                      int nprocs, int direction, void* stream, void* user);
int accfft_amd_plan_set_exchange(accfft_amd_plan* plan, accfft_amd_exchange_fn fn,
                                 void* user);
/* One rank's share of accfft_plan_dft_3d_r2c_gpu(f) on a p0 x p1 grid at
 * coordinates (c0, c1) (src/accfft_gpu.cpp:77-442, the pencil branch
 * :354-375; execute :471-539 forward, :540-623 inverse): n is the global
 * size; input [isize0][isize1][n2] real, output [n0][osize1][osize2]
 * complex, sizes and starts as accfft_local_size_dft_r2c_gpu
 * (src/accfft_common.cpp:98-192, ceil-block splits :49-91, any n and any
 * grid: ragged last blocks and empty ranks included).  execute_r2c runs z
 * r2c; T_plan_1 on the row communicator (the p1 ranks of this c0:
 * memcpy_v1_h1 pack of the z column ranges, exchange, memcpy_v1_h2 unpack,
 * src/transpose_gpu.cpp:1803-2130) to [isize0][n1][osize2]; y c2c; T_plan_2
 * on the column communicator (the p0 ranks of this c1: the y row ranges
 * packed, exchanged, received straight into the output rows of each
 * sender's x range); x c2c.  execute_c2r runs the inverse (x c2c in place
 * on its input, which it overwrites).  The exchanges go through the
 * plan's exchange_v hook; with none each exchange is the identity (the
 * packed blocks taken as received), allowed only where every block sent
 * has the size of the block received from that peer (else execute
 * returns ACCFFT_AMD_ERR_UNSUPPORTED).  When n1 splits over p0 and over
 * p1 in equal power-of-two blocks (or a communicator has one rank) the
 * packs and unpacks are the z and y passes' own stores and loads; other
 * splits run them as separate copies.  p0, p1 <= 64; out of place; no
 * split-precision flags (ACCFFT_MEASURE is accepted: the passes use the
 * built-in tile rule, or accfft_amd_plan_set_tuning); lines must fit one
 * LDS tile (FP32 <= 10240, FP64 <= 5120 points, powers of two up to 4096
 * in FP64); full xyz only. */
int accfft_amd_plan_dft_3d_r2c_pencil(const int n[3], int p0, int p1, int c0, int c1,
                                      void* data, void* data_out, int fp, unsigned flags,
                                      accfft_amd_plan** plan);
/* Exchange hook of a pencil plan: the MPI_Alltoallv of the reference's
 * pencil transposes (src/transpose_gpu.cpp:2058, :1701) on communicator
 * comm (0 = row: the p1 ranks sharing c0, ordered by c1, T_plan_1 / _1i;
 * 1 = column: the p0 ranks sharing c1, ordered by c0, T_plan_2 / _2i).
 * send_bytes[q] bytes at send + send_offsets[q] go to peer q;
 * recv_bytes[q] bytes from peer q land at recv + recv_offsets[q] (host
 * arrays of nprocs entries, device buffers), on the given stream (a
 * hipStream_t); direction -1 in execute_r2c, +1 in execute_c2r.  Called on
 * every rank of the communicator, also where this rank's share is empty.
 * Returns a status.  A RCCL send/recv group over device buffers is the
 * intended implementation. */
typedef int (*accfft_amd_exchange_v_fn)(int comm, const void* send, const int64_t* send_bytes,
                                        const int64_t* send_offsets, void* recv,
                                        const int64_t* recv_bytes, const int64_t* recv_offsets,
                                        int nprocs, int direction, void* stream, void* user);
int accfft_amd_plan_set_exchange_v(accfft_amd_plan* plan, accfft_amd_exchange_v_fn fn,
                                   void* user);
/* accfft_destroy_plan_gpu(f) (src/accfft_gpu.cpp:1150-1192).  Also frees the
 * plan object itself (the reference leaks it). */
int accfft_amd_destroy_plan(accfft_amd_plan* plan);
/* Stream for all later calls on this plan (a hipStream_t, NULL = null stream). */
int accfft_amd_plan_set_stream(accfft_amd_plan* plan, void* stream);
/* Plan facts the drop-in wrappers mirror into accfft_plan_gpu(f):
 * n, isize, istart, osize, ostart (3 each), alloc_max, fp, kind (1 = r2c,
 * 2 = c2c, 3 = r2c slab share, 4 = r2c pencil share), inplace. */
int accfft_amd_plan_info(const accfft_amd_plan* plan, int n[3], int isize[3],
                         int istart[3], int osize[3], int ostart[3],
                         int64_t* alloc_max, int* fp, int* kind, int* inplace);
/* The tile configuration of each axis's strided C2C passes (no counterpart
 * in the reference; what its MEASURE planning would record): -1 the
 * built-in rule, -2 one-level tiles, v >= 0 two-level variant v.  get:
 * what ACCFFT_MEASURE chose; set: force one (ERR_INVALID when the axis
 * length has no such variant), e.g. to test each configuration.
 * On an FP32 r2c plan whose z pass carries the first radix-8 level of the y
 * transform (N1 = 8 M, M = 16 ... 256, N2 = 256 ... 2048, not split),
 * get reports -3 on the y axis when that fusion is off; set(1, -3) turns
 * it off (separate z and N1-point y passes, tuned as above) and
 * set(1, -1) restores it. */
int accfft_amd_plan_get_tuning(const accfft_amd_plan* plan, int choice[3]);
int accfft_amd_plan_set_tuning(accfft_amd_plan* plan, int axis, int choice);
/* Default data pointers recorded at plan time (src/accfft_gpu.cpp:447-450). */
int accfft_amd_plan_data(const accfft_amd_plan* plan, void** data, void** data_out);

/* ---------------------------------------------------------------- execute */
/* accfft_execute_r2c_gpu(f) (src/accfft_gpu.cpp:635-653) and
 * accfft_execute_gpu(dir=-1) (:444-539).  xyz bit d selects axis d (bit 0 =
 * x), as std::bitset<3>.  NULL pointers take the plan's pointers.  timer[4]
 * accumulates the local FFT time in seconds and timer[2] the time of a slab
 * plan's exchange hook (communication), as doc/timing.md; every other call
 * that takes a timer adds its time to timer[4]. */
int accfft_amd_execute_r2c(accfft_amd_plan* plan, const void* in, void* out,
                           double* timer, unsigned xyz);
/* accfft_execute_c2r_gpu(f) (:655-675), accfft_execute_gpu(dir=+1)
 * (:540-603).  Overwrites its input unless ACCFFT_AMD_PRESERVE_INPUT. */
int accfft_amd_execute_c2r(accfft_amd_plan* plan, void* in, void* out,
                           double* timer, unsigned xyz);
/* accfft_execute_c2c_gpu(f) (:943-1136). */
int accfft_amd_execute_c2c(accfft_amd_plan* plan, int direction, void* in,
                           void* out, double* timer, unsigned xyz);
/* Per-axis transforms accfft_execute_{r2c,c2r}_{z,y,x}_gpu_t
 * (:1212-1604).  axis 2 (z): real [N0][N1][N2] <-> complex [N0][N1][N2/2+1];
 * axis 1 (y): <-> [N0][N1/2+1][N2]; axis 0 (x): <-> [N0/2+1][N1][N2].  The
 * input is never overwritten. */
int accfft_amd_execute_r2c_axis(accfft_amd_plan* plan, int axis, const void* in,
                                void* out, double* timer);
int accfft_amd_execute_c2r_axis(accfft_amd_plan* plan, int axis, const void* in,
                                void* out, double* timer);

/* ---------------------------------------------------------------- operators */
/* accfft_grad_gpu(f) (src/operators_gpu.cpp:85-88; operators_gpu.txx:172-278):
 * component d = c2r_d(i*k_d/N_d * r2c_d(A)), one fused pass per component.
 * mask bit d selects the component (pXYZ); NULL outputs are skipped.
 * slow != 0 gives accfft_grad_gpu_slow_t (:72-170). */
int accfft_amd_grad(accfft_amd_plan* plan, void* Ax, void* Ay, void* Az,
                    const void* A, unsigned mask, double* timer, int slow);
/* accfft_laplace_gpu(f) (operators_gpu.txx:280-334). */
int accfft_amd_laplace(accfft_amd_plan* plan, void* LA, const void* A, double* timer);
/* accfft_biharmonic_gpu(f) (operators_gpu.txx:555-609). */
int accfft_amd_biharmonic(accfft_amd_plan* plan, void* BA, const void* A, double* timer);
/* Inverse Laplacian: CPU-only upstream (src/operators.txx:276-333,
 * :1270-1321); here on the GPU. */
int accfft_amd_inv_laplace(accfft_amd_plan* plan, void* out, const void* A, double* timer);
/* accfft_amd_inv_laplace without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight solver loops, harnesses that forbid
 * hipMalloc).  Same kernels and result as accfft_amd_inv_laplace, but the
 * intermediate spectrum lives in a fixed file-scope device array of
 * ACCFFT_AMD_NOALLOC_MAX_BYTES per GPU instead of the plan's workspace;
 * plans whose spectrum is larger get ACCFFT_AMD_ERR_SCRATCH, plans with an
 * axis longer than one LDS tile (whose passes need workspaces of their
 * own) ACCFFT_AMD_ERR_UNSUPPORTED.  The array is
 * shared by every call on a GPU (and keeps the last call's contents), so
 * calls that overlap on one GPU -- on different streams or host threads --
 * race; order them on one stream.  The current device must be the plan's.
 * Per-launch profiling is not recorded for this call. */
int accfft_amd_inv_laplace_noalloc(accfft_amd_plan* plan, void* out, const void* A,
                                   double* timer);
/* Inverse biharmonic: CPU-only upstream (src/operators.txx:336-392). */
int accfft_amd_inv_biharmonic(accfft_amd_plan* plan, void* out, const void* A, double* timer);
/* FFT Poisson solve of the step4 program (steps/step4/step4_gpu.cpp:520-640:
 * accfft_execute_r2c_gpu, poisson_fourier_filter_gpu, accfft_execute_c2r_gpu)
 * on an r2c plan, with the filter fused into the x pass: phi = the
 * zero-mean solution of laplace(phi) = rho on the unit box.  method 0: the
 * discrete 7-point Laplacian's eigenvalues; 1: -4 pi^2 |k|^2
 * (steps/step4/kernels.cu:24-110).  rho is not modified. */
int accfft_amd_poisson_solve(accfft_amd_plan* plan, int method, void* phi, const void* rho,
                             double* timer);
/* accfft_divergence_gpu(f) (operators_gpu.txx:448-553); slow != 0 gives
 * accfft_divergence_gpu_slow_t (:336-446). */
int accfft_amd_divergence(accfft_amd_plan* plan, void* divA, const void* Ax,
                          const void* Ay, const void* Az, double* timer, int slow);

/* Element-wise spectral multiply, the extern "C" multipliers of
 * src/operators_cuda.cu:421-486 (grad_mult_wave_number{x,y,z}_gpu_c(f),
 * laplace_mult_wave_number_gpu_c(f), biharmonic_mult_wave_number_gpu_c(f))
 * plus the inverse symbols.  kind: 0/1/2 grad x/y/z, 3 laplace,
 * 4 biharmonic, 5 inverse laplace, 6 inverse biharmonic, 7 / 8 the Poisson
 * filter of methods 0 / 1 (accfft_amd_poisson_filter).  scale_xyz selects
 * the axes whose lengths divide the result (the grad launchers' bitset,
 * operators_cuda.cu:219-229; the other kinds use all three). */
int accfft_amd_mult_wave_number(int kind, int fp, void* wA, const void* A,
                                const int n[3], const int osize[3],
                                const int ostart[3], unsigned scale_xyz,
                                void* stream);
/* poisson_fourier_filter_gpu (steps/step4/kernels.cu:24-146): in place on a
 * half spectrum block osize at offset ostart of an n[0] x n[1] x n[2] r2c
 * transform, every mode but (0,0,0) divided by its Poisson eigenvalue times
 * n[0] n[1] n[2], the mean set to zero.  method 0 or 1 as
 * accfft_amd_poisson_solve.  Stream-ordered. */
int accfft_amd_poisson_filter(int fp, void* data_hat, const int n[3], const int osize[3],
                              const int ostart[3], int method, void* stream);
/* daxpy_gpu_c(f) (operators_cuda.cu:470-474): y += alpha * x. */
int accfft_amd_axpy(int fp, int64_t n, double alpha, const void* x, void* y,
                    void* stream);

/* ---------------------------------------------------------------- adjoints */
/* Adjoints of the r2c plan's real-linear maps (no counterpart in the
 * reference), with <a, b> = Re sum conj(a) b on half spectra and the dot
 * product on real fields.  None overwrites its input.
 *   r2c^T(G) = c2r(w G), w = 1/2 on z bins other than DC / Nyquist;
 *   c2r^T(y) = r2c(y) / w;
 *   grad^T(gx, gy, gz) = -div (NULL components are skipped);
 *   div^T(v) = -grad(v) (NULL outputs are skipped);
 * laplace, biharmonic and the inverse laplace / biharmonic are self-adjoint
 * (their adjoint is the forward call). */
int accfft_amd_adjoint_r2c(accfft_amd_plan* plan, const void* G, void* x, double* timer);
int accfft_amd_adjoint_c2r(accfft_amd_plan* plan, const void* y, void* G, double* timer);
int accfft_amd_adjoint_grad(accfft_amd_plan* plan, void* out, const void* gx,
                            const void* gy, const void* gz, double* timer);
int accfft_amd_adjoint_divergence(accfft_amd_plan* plan, void* ox, void* oy, void* oz,
                                  const void* v, double* timer);
/* Frees the plan's workspaces (intermediate spectra of the operators); they
 * are reallocated on next use.  Lets a caller trade recomputation for memory. */
int accfft_amd_plan_release_workspace(accfft_amd_plan* plan);
int accfft_amd_plan_workspace_bytes(const accfft_amd_plan* plan, int64_t* bytes);

/* ---------------------------------------------------------------- transposes */
/* Local (on-GPU) transposes of src/transpose_cuda.cu, bit-exact
 * permutations of tuples of elem_size-byte elements.
 * local_transpose_cuda(r,c,t,in,out) (:242-257, kernel cuda1 :6-11):
 *   [r][c][t] -> [c][r][t]. */
int accfft_amd_local_transpose(int r, int c, int64_t t, const void* in, void* out,
                               int elem_size, void* stream);
/* local_transpose_cuda(r,c,t,t2,in,out) (:207-224, cuda2 :13-21): unpack with
 * a ragged last row of t2-element tuples. */
int accfft_amd_local_transpose_ragged(int r, int c, int64_t t, int64_t t2,
                                      const void* in, void* out, int elem_size,
                                      void* stream);
/* local_transpose_col_cuda(r,c,t,t2,in,out) (:78-95, cuda4 :59-66): pack
 * with a ragged last column. */
int accfft_amd_local_transpose_col(int r, int c, int64_t t, int64_t t2,
                                   const void* in, void* out, int elem_size,
                                   void* stream);
/* local_transpose_cuda(r,c,t,A) (:227-239, cuda3 :23-57): in place,
 * cycle-following. */
int accfft_amd_local_transpose_inplace(int r, int c, int64_t t, void* A,
                                       int elem_size, void* stream);
/* memcpy_v1_h1 (src/transpose_cuda.cu:97-143, called by the pencil
 * transposes at src/transpose_gpu.cpp:1888, :2284): pack for an exchange
 * over nprocs_1 ranks.  For each rank p, batch h < howmany and row
 * i < local_n0, the local_n1_proc[p] tuples of n_tuples elements that start
 * at column local_1_start_proc[p] of row i (data[h * idist + (i * N1 +
 * start) * n_tuples]) go to the next place of send_recv, in the order
 * (p, h, i).  The per-rank arrays are host memory; nprocs_1 <= 64.
 * Bit-exact copy of elem_size-byte elements, stream-ordered. */
int accfft_amd_pencil_pack(int nprocs_1, int howmany, int local_n0, int64_t n_tuples,
                           const int64_t* local_n1_proc, void* send_recv, const void* data,
                           int64_t idist, int64_t N1, const int64_t* local_1_start_proc,
                           int elem_size, void* stream);
/* memcpy_v1_h2 (src/transpose_cuda.cu:145-197, called at
 * src/transpose_gpu.cpp:2118, :2414): unpack after the exchange over
 * nprocs_0 ranks.  send_recv holds segments of local_n1 * n_tuples elements
 * in the order (p, h, i) for i in [local_0_start_proc[p],
 * local_0_start_proc[p] + local_n0_proc[p]); each goes to data[h * odist +
 * i * local_n1 * n_tuples].  Host per-rank arrays; nprocs_0 <= 64. */
int accfft_amd_pencil_unpack(int nprocs_0, int howmany, const int64_t* local_0_start_proc,
                             const int64_t* local_n0_proc, void* data, int64_t odist,
                             int64_t local_n1, int64_t n_tuples, const void* send_recv,
                             int elem_size, void* stream);

/* ---------------------------------------------------------------- misc */
/* Synchronise the plan's stream (what the reference's blocking calls do at
 * src/accfft_gpu.cpp:620). */
int accfft_amd_plan_sync(accfft_amd_plan* plan);
/* Time of the last kernel launched by this plan's most recent call with a
 * timer, for benchmarking: device milliseconds of [first, last] launch. */
int accfft_amd_last_elapsed_ms(const accfft_amd_plan* plan, float* ms);

/* Self-test of the matrix-core radix-16 DFT (no counterpart in the
 * reference): one wavefront transforms 32 vectors of 16 complex FP32 values
 * (x, y: 512 complex each, vector-major) with split kind 1 (BF16), 2
 * (BF16X3) or 3 (FP16) and sign -1 / +1.  Pins the matrix-core lane layout
 * and the split error in the tests. */
int accfft_amd_selftest_dft16(int split, int sign, const void* x, void* y, void* stream);

/* Per-launch profiling (no counterpart in the reference, whose only
 * instrumentation is the timer[] array, doc/timing.md).  When enabled, every
 * kernel launch of the plan is bracketed by HIP events on the plan's stream;
 * totals are grouped by kernel kind, named "<z|s>fft_<mode>_n<N>_<axis>_<fp>".
 * accfft_amd_profile_get synchronises on the recorded events. */
int accfft_amd_plan_set_profiling(accfft_amd_plan* plan, int enable);
int accfft_amd_profile_reset(accfft_amd_plan* plan);
int accfft_amd_profile_count(accfft_amd_plan* plan, int* nkinds);
int accfft_amd_profile_get(accfft_amd_plan* plan, int idx, char* name, int name_len,
                           double* total_ms, int64_t* launches, double* bytes_per_launch);

#ifdef __cplusplus
}
#endif

#endif /* ACCFFT_AMD_H */
