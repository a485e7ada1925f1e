/* kutta_hip.h — C ABI of the MI355X (gfx950) HIP backend for Kutta's SIMPLEC
 * pressure-correction path.
 *
 * This is the drop-in seam behind Kutta's abstract device runtime
 * (src/kutta/runtime/device.hpp:107-240). A HipDevice : kutta::Device adapter
 * forwards each virtual to one entry point below; buffers are plain device
 * pointers and scalars travel by value, so no C++ or HIP type crosses the ABI.
 *
 * Conventions:
 *   - Every function returning int returns KH_OK (0) on success. On failure it
 *     returns KH_ERROR (or KH_NONFINITE) and kh_last_error() describes it.
 *   - All work is ordered on one in-order stream per device, in call order
 *     (device.hpp:7-9). Host-visible results (download, host reductions) are
 *     complete on return; everything else is asynchronous.
 */
#ifndef KUTTA_HIP_H
#define KUTTA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { KH_OK = 0, KH_ERROR = 1, KH_NONFINITE = 2 };

typedef struct kh_device kh_device;

/* Replaces CudaDevice construction / Device::create(Backend::Cuda)
 * (runtime/device.cpp:32-49, runtime/cuda/cuda_device.cu:112-131). */
int kh_create(int ordinal, kh_device** out);
/* Replaces ~CudaDevice (cuda_device.cu:121-124). */
void kh_destroy(kh_device* dev);
/* Device::name() (device.hpp:117). */
const char* kh_name(const kh_device* dev);
/* Message for the most recent failure on this thread. */
const char* kh_last_error(void);
/* Number of visible devices (Device::enumerate, device.cpp:60-88). */
int kh_device_count(void);

/* Device::alloc_bytes / BufferHandle lifetime (device.hpp:234; cuda_device.cu:17-29).
 * A zero-byte allocation returns a unique non-null pointer. */
int kh_alloc(kh_device* dev, size_t bytes, void** out);
int kh_free(kh_device* dev, void* ptr);
/* Device::upload_bytes: returns once host memory is reusable (device.hpp:235). */
int kh_upload(kh_device* dev, void* dst, const void* src, size_t bytes);
/* Device::download_bytes: complete on return (device.hpp:236). */
int kh_download(kh_device* dev, const void* src, void* dst, size_t bytes);
/* Device::copy_bytes: ordered, asynchronous; src == dst is a no-op (device.hpp:237). */
int kh_copy(kh_device* dev, void* dst, const void* src, size_t bytes);
/* Device::fill_bytes with 32-bit words (device.hpp:238; cuda_device.cu:204-213).
 * Implemented as a kernel, so it may be recorded into a graph. */
int kh_fill32(kh_device* dev, void* dst, uint32_t word, size_t count);
/* Device::synchronize (device.hpp:208). */
int kh_synchronize(kh_device* dev);

/* One kernel argument (KernelArg, device.hpp:73-102): a device buffer or a
 * scalar of 4 bytes (float/int32/uint32) passed by value. */
enum { KH_ARG_BUFFER = 0, KH_ARG_SCALAR = 1 };
typedef struct kh_arg {
  uint32_t kind;
  uint32_t size;
  union {
    void* ptr;
    float f32;
    int32_t i32;
    uint32_t u32;
    uint64_t bits;
  } value;
} kh_arg;

/* Kernel ids. The first block names the reference kernels on the path; their
 * parameter order is the reference .kern.h signature order (kernelgen.py:150-160).
 * The KH_FUSED_* block are MI355X-specific fused kernels with no reference
 * counterpart; each is bitwise equal to the reference sequence it names. */
typedef enum kh_kernel {
  KH_KERNEL_NONE = 0,
  /* kernels/linalg.kern.h */
  KH_CSR_SPMV = 1,               /* :5-16   rows,cols,values,x,y */
  KH_CSR_SPMV_GUARDED,           /* :111-114 rows,cols,values,x,y,status */
  KH_CSR_RESIDUAL,               /* :47-50  rows,cols,values,x,b,r */
  KH_CSR_RESIDUAL_GUARDED,       /* :115-118 rows,cols,values,x,b,r,status */
  KH_KRYLOV_DIAGONAL,            /* :52-70  rows,diag,values,inverse,invalid,use_l1 */
  KH_KRYLOV_MULTIPLY,            /* :72-75  r,inverse,z */
  KH_KRYLOV_MULTIPLY_GUARDED,    /* :119-122 r,inverse,z,status */
  KH_KRYLOV_SMOOTH,              /* :77-80  r,inverse,x,omega */
  KH_KRYLOV_SHIFT,               /* :82     x,mean */
  KH_KRYLOV_DIRECTION,           /* :84-86  z,p,beta */
  KH_KRYLOV_STEP,                /* :88-92  x,r,p,ap,alpha */
  KH_KRYLOV_BICG_DIRECTION,      /* :94-97  r,v,p,beta,omega */
  KH_KRYLOV_DIFFERENCE,          /* :99-102 a,b,out,scale */
  KH_KRYLOV_BICG_STEP,           /* :104-107 x,p,s,alpha,omega */
  KH_AMG_COARSE_SOLVE,           /* :126-147 rows,cols,values,x,b,inverse,status,count,sweeps */
  KH_AMG_SMOOTH,                 /* :150-156 rows,cols,values,x,b,inverse,next,status */
  KH_AMG_SMOOTH_ZERO,            /* :162-170 rows,cols,values,x,b,inverse,next,status */
  /* kernels/amg.kern.h */
  KH_AMG_RESTRICT,               /* :5-16   offsets,members,fine,coarse */
  KH_AMG_PROLONG,                /* :18-21  aggregate,coarse,fine */
  KH_AMG_GALERKIN,               /* :27-38  offsets,entries,fine,coarse */
  KH_AMG_M_MATRIX_DIAGONAL,      /* :44-58  offsets,members,fine_rows,fine_diag,fine,rows,diag,coarse */
  KH_AMG_RESTRICT_GUARDED,       /* :84-87  offsets,members,fine,coarse,status */
  KH_AMG_PROLONG_GUARDED,        /* :88-91  aggregate,coarse,fine,status */
  /* kernels/blas.kern.h */
  KH_AXPY,                       /* :5  y,x,a */
  KH_XPBY,                       /* :8  y,x,b */
  KH_AXPBYZ,                     /* :11 z,x,y,a,b */
  KH_SCALE_COPY,                 /* :16 y,x,a */
  KH_MUL,                        /* :19 z,x,y */
  /* kernels/scalars.kern.h */
  KH_SOLVER_LIMITS,              /* :51-55  scalars,tol_hi,tol_lo,div_hi,div_lo */
  KH_SOLVER_RESET,               /* :56-62  scalars,status,iterations */
  KH_SOLVER_PCG_ALPHA,           /* :63-71  scalars,status */
  KH_SOLVER_PCG_BETA,            /* :72-80  scalars,status,flexible */
  KH_SOLVER_BICG_BETA,           /* :81-89  scalars,status */
  KH_SOLVER_BICG_ALPHA,          /* :90-97  scalars,status */
  KH_SOLVER_BICG_S,              /* :98-103 scalars,status */
  KH_SOLVER_BICG_OMEGA,          /* :104-112 scalars,status */
  KH_SOLVER_CHECK,               /* :113-121 scalars,status,pcg */
  KH_SOLVER_STEP,                /* :122-127 x,r,p,q,scalars,status */
  KH_SOLVER_DIRECTION,           /* :128-131 z,p,scalars,status */
  KH_SOLVER_BICG_DIRECTION,      /* :132-137 r,v,p,scalars,status */
  KH_SOLVER_BICG_DIFFERENCE,     /* :138-141 r,q,s,scalars,status */
  KH_SOLVER_BICG_STEP,           /* :142-149 x,r,z,zs,s,t,scalars,status */
  KH_SOLVER_PROJECT,             /* :150-152 x,sum,count_hi,count_lo */

  /* kernels/fv_flow.kern.h (finite-volume face loops and assembly); arguments
     in the reference's parameter order. */
  KH_FV_PRESSURE_ASSEMBLE = 100, /* :164-183 offsets,faces,sign,adjacent,entry,rows,diag,
                                    coefficient,flux,values,rhs,reference */
  KH_FV_VELOCITY_BOUNDARY,       /* :5-32   n = boundary faces */
  KH_FV_PRESSURE_BOUNDARY,       /* :34-41  n = boundary faces */
  KH_FV_PRESSURE_FACES,          /* :43-65  n = faces */
  KH_FV_MOMENTUM_SOURCE,         /* :67-78  n = cells */
  KH_FV_MOMENTUM_FACES,          /* :83-109 n = faces */
  KH_FV_CONVECTIVE_OUTFLOW,      /* :111-116 n = cells */
  KH_FV_PECLET_SIDE_VELOCITY,    /* :122-133 n = internal faces */
  KH_FV_RHIE_CHOW,               /* :139-162 n = faces */
  KH_FV_CORRECT_FLUX,            /* :185-190 n = faces */
  KH_FV_CORRECT_VELOCITY,        /* :192-202 n = cells */
  KH_FV_MONITOR_CELL,            /* :204-215 n = cells */
  KH_FV_SPEED,                   /* :217-219 n = cells */
  KH_FV_CONTINUITY,              /* :221-226 n = cells */
  KH_FV_WALL_PRESSURE_FORCE,     /* :228-233 n = boundary faces */
  KH_FV_WALL_VISCOUS_FORCE,      /* :235-257 n = boundary faces */
  KH_FV_FLUENT_MOMENTUM_SCALE,   /* :261-266 n = cells */
  KH_FV_PHYSICAL_TIME_TERMS,     /* :268-276 n = cells */
  KH_FV_FACE_TIME_DEFECT,        /* :278-281 n = faces */
  KH_FV_FACE_TIME_FLUX,          /* :283-295 n = faces */
  KH_FV_CELL_MASS_ERROR,         /* :297-299 n = cells */
  /* kernels/fv_gradient.kern.h (least-squares gradients and limiters) */
  KH_FV_GRADIENT,                /* :6-23 */
  KH_FV_PAIRED_GRADIENT,         /* :28-56 */
  KH_FV_RANGE_SHIFT,             /* :72-74 */
  KH_FV_LIMITER,                 /* :76-94 */
  KH_FV_PAIRED_LIMITER,          /* :103-123 */
  KH_FV_PAIRED_GRADIENT_LIMITER, /* :127-145 */
  KH_FV_VECTOR_GRADIENT,         /* :147-166 */
  KH_FV_SCALE_GRADIENT,          /* :168-171 */
  KH_FV_LIMITER_SCALE,           /* :173-180 */
  KH_FV_PAIRED_LIMITER_SCALE,    /* :182-189 */
  KH_FV_PAIRED_GRADIENT_LIMITER_SCALE,/* :191-199 */
  KH_FV_JUMP_GRADIENT,           /* :202-221 */
  KH_FV_FACE_JUMP,               /* :227-240 */
  KH_FV_GATHER_JUMP_GRADIENT,    /* :242-254 */
  KH_FV_PAIRED_JUMP_GRADIENT,    /* :258-292 */
  KH_FV_LSQ_RESTORE_INVERSE,     /* :296-299 */
  /* kernels/fv_transport.kern.h (transport faces and matrix assembly) */
  KH_FV_DIFFUSION_CORRECTION_BOUND,/* :3-10 */
  KH_FV_TRANSPORT_FACES,         /* :12-35 */
  KH_FV_TRANSPORT_ASSEMBLE,      /* :37-112 */
  KH_FV_FIXED_ROWS,              /* :115-124 */
  KH_FV_FLOOR,                   /* :126-128 */
  KH_FV_SCALAR_STEADY_MEASURE,   /* :130-135 */
  KH_FV_SCALAR_SET_FIXED,        /* :137-139 */
  KH_FV_SCALAR_MONITOR,          /* :141-148 */
  /* kernels/fv_paired.kern.h (float-pair SIMPLEC state) */
  KH_FV_COMPENSATED_PRESSURE_JUMP,   /* :5-33 */
  KH_FV_PAIRED_PRESSURE_JUMP,        /* :46-65 */
  KH_FV_COMPENSATED_RHIE_CHOW,       /* :69-85 */
  KH_FV_PAIRED_PRESSURE_BOUNDARY_LOW,/* :89-97 */
  KH_FV_PAIRED_PRESSURE_FACES,       /* :98-128 */
  KH_FV_PAIRED_MOMENTUM_RESIDUAL,    /* :140-152 */
  KH_FV_COMPENSATED_MOMENTUM_RESIDUAL,/* :156-170 */
  KH_FV_PAIRED_ACCUMULATE,           /* :172-174 */
  KH_FV_COMPENSATED_PRESSURE_FORCE,  /* :179-196 */
  KH_FV_PAIRED_VELOCITY_BOUNDARY,    /* :200-218 */
  KH_FV_PAIRED_ADVECTION,            /* :220-236 */
  KH_FV_PAIRED_FACE_RESIDUAL,        /* :238-247 */
  KH_FV_PAIRED_FLUX,                 /* :259-290 */
  KH_FV_PAIRED_FLUX_OFFSET,          /* :292-303 */
  KH_FV_PAIRED_CONTINUITY,           /* :308-318 */
  KH_FV_PAIRED_PRESSURE_RHS,         /* :321-323 */
  KH_FV_PAIRED_CORRECT_FLUX,         /* :326-333 */
  KH_FV_PAIRED_CORRECT_VELOCITY,     /* :336-347 */
  KH_FV_PAIRED_VALUE,                /* :349-351 */
  KH_FV_PAIRED_VECTOR_VALUES,        /* :352-355 */
  KH_FV_COMPENSATED_CORRECT_FLUX,    /* :356-363 */
  KH_FV_COMPENSATED_CORRECT_VELOCITY,/* :365-377 */
  /* kernels/fv_rans.kern.h (turbulent stresses, momentum boundary terms) */
  KH_FV_BACKFLOW_PRESSURE,           /* :4-10 */
  KH_FV_EFFECTIVE_VISCOSITY,         /* :12-17 */
  KH_FV_TRANSPOSE_STRESS,            /* :21-37 */
  KH_FV_STRESS_SOURCE,               /* :38-46 */
  KH_FV_BOUNDARY_MOMENTUM,           /* :47-72 */
  KH_FV_MOMENTUM_DIAGONAL,           /* :75-78 */
  KH_FV_SST_FORCE,                   /* :79-83 */
  KH_FV_TURBULENCE_BOUNDARY,         /* :86-101 */
  /* kernels/sst.kern.h (k-omega SST; tanh/log/pow from the device library) */
  KH_SST_RATES,                      /* :40-50 */
  KH_SST_BLENDING,                   /* :52-70 */
  KH_SST_VISCOSITY,                  /* :72-88 */
  KH_SST_BLENDING_VISCOSITY,         /* :90-103 */
  KH_SST_WALL,                       /* :105-151 */
  KH_SST_WALL_OMEGA,                 /* :153-169 */
  KH_SST_WALL_CELLS,                 /* :171-190 */
  KH_SST_SOURCES,                    /* :192-236 */
  KH_SST_FACE_DIFFUSION,             /* :238-253 */
  KH_SST_SUSTAIN,                    /* :255-259 */
  /* kernels/fv_fluent_schemes.kern.h */
  KH_FV_SIMPLE_INVERSE,              /* :10-12 */
  KH_FV_STANDARD_PRESSURE_FACES,     /* :28-39 */
  KH_FV_CELL_PRESSURE_BOUNDARY,      /* :47-50 */
  KH_FV_FORCE_TO_SHEAR,              /* :54-59 */
  KH_FV_POOR_CELL_PRESSURE_FACES,    /* :70-84 */
  KH_FV_POOR_PRESSURE_INTERFACES,    /* :93-97 */

  /* Fused MI355X kernels (no reference counterpart). */
  KH_FUSED_SPMV_DOT_ALPHA = 200, /* csr_spmv_guarded + dot_into(p,q,2) + solver_pcg_alpha:
                                    rows,cols,values,p,q,scalars,status */
  KH_FUSED_STEP_DOT_CHECK,       /* solver_step + dot_into(r,r,18) + solver_check(pcg=1):
                                    x,r,p,q,scalars,status */
  KH_FUSED_RESIDUAL_RESTRICT,    /* csr_residual_guarded + amg_restrict_guarded (fine r unused):
                                    rows,cols,values,x,b,offsets,members,coarse_b,status */
  KH_FUSED_PROLONG_SMOOTH,       /* amg_prolong_guarded + amg_smooth (x + P c, then sweep):
                                    rows,cols,values,x,b,inverse,next,aggregate,coarse,status */
  KH_FUSED_SMOOTH_DOT_BETA,      /* amg_smooth + dot_into(r_outer,next,4) + solver_pcg_beta(0):
                                    rows,cols,values,x,b,inverse,next,status,scalars */
  KH_FUSED_SMOOTH_SUM,           /* amg_smooth + sum_into(next -> pair at sum[0..1]):
                                    rows,cols,values,x,b,inverse,next,status,sum */
  KH_FUSED_PROJECT_DOT_BETA,     /* solver_project(z) + dot_into(r,z,4) + solver_pcg_beta(0):
                                    z,sum,count_hi,count_lo,r,scalars,status */
  KH_FUSED_SMOOTH_FROM_ZERO,     /* fill(x,0) + amg_smooth_zero, without storing the zeros:
                                    rows,cols,values,b,inverse,next,status */
  KH_FUSED_COARSE_TAIL,          /* the whole V-cycle (linalg/amg.cpp:283-310) from one level
                                    down to the coarsest in one workgroup, iterates in LDS:
                                    levels,count,pre,post,coarse_sweeps,nullspace,b,x,status
                                    (n = 1; levels is a device array of kh_tail_level) */
  /* BiCGStab-Jacobi (linalg/solver.cpp:314-361), one iteration in five launches: */
  KH_FUSED_DOT_BICG_BETA,        /* dot_into(shadow,r,0) + solver_bicg_beta: shadow,r,scalars,status */
  KH_FUSED_BICG_DIRECTION,       /* solver_bicg_direction + krylov_multiply_guarded(p -> z):
                                    r,q,p,inverse,z,scalars,status */
  KH_FUSED_SPMV_SHADOW_ALPHA,    /* csr_spmv_guarded(z -> q) + dot_into(shadow,q,2) + solver_bicg_alpha:
                                    rows,cols,values,z,q,shadow,scalars,status */
  KH_FUSED_BICG_DIFFERENCE_S,    /* solver_bicg_difference + dot_into(s,s,16) + solver_bicg_s +
                                    krylov_multiply_guarded(s -> zs): r,q,s,inverse,zs,scalars,status */
  KH_FUSED_SPMV_DOT2_OMEGA,      /* csr_spmv_guarded(zs -> t) + dot2_into(t,t,t,s,2) + solver_bicg_omega:
                                    rows,cols,values,zs,t,s,scalars,status */
  KH_FUSED_BICG_STEP_CHECK_BETA, /* solver_bicg_step + dot_into(r,r,18) + solver_check(pcg=0), then the
                                    next iteration's dot_into(shadow,r,0) + solver_bicg_beta:
                                    x,r,z,zs,s,t,shadow,scalars,status */
  /* Flexible GMRES(m) (no reference counterpart; Saad 1993): basis vectors
     FP32 (column-major, n per vector), Hessenberg, rotations and residual
     recurrence FP64 in a device block of (m+1)m + 5m + 3 doubles; partials
     hold 64 doubles per 256-row workgroup. */
  KH_GMRES_START,                /* beta = ||r||, g = beta e1, V0 = r/beta: r,V,block,m,status,partials */
  KH_GMRES_ARNOLDI,              /* CGS2 of w against V_0..V_j, rotation, estimate check, V_{j+1}:
                                    V,w,block,m,j,status,partials */
  KH_GMRES_UPDATE,               /* y = H^-1 g over status[2] columns, x += Z y: Z,x,block,m,status */
  /* FP64 dense coarsest-level solve of the mixed-precision AMG path (no
     reference counterpart); the operator must be symmetric positive definite. */
  KH_DENSE_INVERT,               /* blocked Gauss-Jordan inverse of an n-row CSR operator, FP64 (the
                                    panel and update products on the FP64 MFMA), no pivoting:
                                    rows,cols,values,inverse (n*n doubles),scratch (4096 doubles);
                                    n <= 8192 */
  KH_DENSE_APPLY,                /* x = inverse b (FP64 accumulation): inverse,b,x,status */
  /* Learned coarse correction (no reference counterpart): a message-passing
     network of hidden width KH_LEARNED_HIDDEN over a coarse level's graph,
     bf16 MFMA layers with FP32 accumulation. Gradient kernels store FP32
     partials per slot (KH_LEARNED_SLOTS rows of `width` floats, one column
     per parameter plus the loss and the target energy) that
     KH_LEARNED_UPDATE sums in FP64. n is the node count unless stated. */
  KH_LEARNED_GRAPH,              /* node features and neighbour weights of an operator:
                                    row_ptr,col,val,features(2n),ws(nnz),wt(nnz),gstats(1),scratch */
  KH_LEARNED_NORM,               /* slot sums of r^2: r,scratch */
  KH_LEARNED_INPUT,              /* features and input layer: r,scratch,features,gstats,theta,x(4n),h0,stats(2) */
  KH_LEARNED_LAYER,              /* h = relu(h_prev Ws + (S h_prev) Wn + b):
                                    row_ptr,col,ws,h_prev,wf,bias,h,m */
  KH_LEARNED_HEAD,               /* e = (h wo + bo) * scale: h,head,stats,e */
  KH_LEARNED_HEAD_LOSS,          /* y, dy, wo/bo/loss/energy partials: h,head,stats,target,y,dy,partials,width */
  KH_LEARNED_BACK,               /* one layer's backward: last,dy,head,p_in,dm_in,row_ptr,col,wt,h,wb,
                                    dz,p_out,dm_out,partials,width */
  KH_LEARNED_WEIGHT_GRAD,        /* [Ws;Wn] partials: h_prev,m,dz,partials,width */
  KH_LEARNED_BACK_INPUT,         /* W0/b0 partials: p_in,dm_in,row_ptr,col,wt,h0,x,partials,width */
  KH_LEARNED_UPDATE,             /* n = parameters + 2: mode (0 pack, 1 gradient, 2 Adam step),
                                    partials,theta,adam_m,adam_v,grad,loss(2: mean squared error,
                                    mean squared normalised target),layers,step,lr,nodes,wf,wb */
  KH_LEARNED_RESTRICT,           /* n = coarse nodes: r = R (b - A z) through composite aggregates:
                                    row_ptr,col,val,b,z,member_offsets,members,r,scratch (fine rows),
                                    fine rows,group (lanes per coarse node: a power of two <= 64) */
  KH_LEARNED_PROLONG,            /* n = fine rows: z += e[map] (map < 0: none): map,e,z */
  /* Batched BiCGStab-Jacobi (no reference counterpart): the five launches of
     a fused iteration for up to KH_BATCH_MAX_LANES independent systems of n
     rows at once (the three momentum components). Arguments: lanes (device
     array of kh_bicg_lane), lane count, mask of the lanes that step. Each
     lane's iterates are bitwise those of its own KH_FUSED_* launches. */
  KH_BATCH_BICG_DIRECTION,
  KH_BATCH_SPMV_SHADOW_ALPHA,
  KH_BATCH_BICG_DIFFERENCE_S,
  KH_BATCH_SPMV_DOT2_OMEGA,
  KH_BATCH_BICG_STEP_CHECK_BETA,
  /* bfloat16 coarse-level values of the mixed-precision V-cycle (no reference
     counterpart): KH_TO_BF16 rounds FP32 values to nearest even (n = entries:
     src,dst); the four smoothing kernels take the same arguments as their
     FP32 counterparts with `values` in bfloat16, widened to FP32 for the
     row arithmetic. */
  KH_TO_BF16,
  KH_BF16_AMG_SMOOTH,            /* as KH_AMG_SMOOTH */
  KH_BF16_SMOOTH_FROM_ZERO,      /* as KH_FUSED_SMOOTH_FROM_ZERO */
  KH_BF16_RESIDUAL_RESTRICT,     /* as KH_FUSED_RESIDUAL_RESTRICT */
  KH_BF16_PROLONG_SMOOTH,        /* as KH_FUSED_PROLONG_SMOOTH */
  /* fv/device_mesh.cpp:18-64, 136-157, 232-245 (WallBvh::distance): n = cells;
     cell centroids (3n doubles), wall triangles (9 doubles each: a, b, c),
     the bounding box of each group of 32 consecutive triangles (6 doubles:
     lo x y z, hi x y z), triangle count, out (n floats) =
     float(sqrt(min squared FP64 distance)), +inf without triangles. */
  KH_WALL_DISTANCE,
  /* The coarsest level's smoothing (linalg/amg.cpp:287-293: coarse_sweeps
     amg_smooth sweeps from zero, the first as KH_FUSED_SMOOTH_FROM_ZERO) in
     one workgroup, the iterate ping-ponging in LDS, for a level too large
     for KH_FUSED_COARSE_TAIL (no reference counterpart; bitwise equal to the
     per-sweep launches): rows,cols,values,b,inverse,x,status,sweeps
     (n = rows, at most KH_COARSE_SWEEPS_MAX_ROWS; x receives the result). */
  KH_FUSED_COARSE_SWEEPS,
  /* KH_FUSED_RESIDUAL_RESTRICT with 128 coarse rows per workgroup instead of
     64, for aggregates of about two members (no reference counterpart; same
     arguments and bits). */
  KH_FUSED_RESIDUAL_RESTRICT_WIDE,
  /* kernels/scalars.kern.h:157-186: the optional K-cycle's two inner Krylov
     corrections of a coarse equation (linalg/amg.cpp:315-336), on the
     level's 12 pair scalars and 1-int status. */
  KH_AMG_INNER_ALPHA,            /* :157-166 scalars,status,parent_status,first */
  KH_AMG_INNER_FIRST,            /* :167-172 b,p,q,x,t,scalars,status (unguarded) */
  KH_AMG_INNER_CHECK,            /* :173-176 scalars,status */
  KH_AMG_INNER_BETA,             /* :177-181 scalars,status */
  KH_AMG_INNER_STEP,             /* :182-186 x,p,scalars,status */
  /* KH_AMG_SMOOTH for levels with rows longer than KH_WIDE_ROW entries: the
     row's entries loaded in chunks of up to 64 instead of 16 (no reference
     counterpart; same arguments, same bits). */
  KH_AMG_SMOOTH_WIDE,
  /* The K-cycle's inner correction in four fused steps (no reference
     counterpart; amg.cpp:320-336 with each dot product in its producer and
     each scalar kernel in the reduction's finisher; n = the level's rows;
     spd 1 for A-conjugate steps, 0 for GCR):
     FIRST_STEP   p = z (unguarded), q = A z, (d.q, b.d) -> scalars 0, 2,
                  amg_inner_alpha(first): rows,cols,values,z,p,q,b,scalars,
                  status,parent_status,spd
     FIRST_UPDATE amg_inner_first, (t.t, b.b) -> scalars 4, 6,
                  amg_inner_check: b,p,q,x,t,scalars,status
     BETA_STEP    w = z or A z, w.q -> scalars 4, amg_inner_beta:
                  rows,cols,values,z,w,q,scalars,status,spd
     SECOND_STEP  q = A p, (d.q, t.d) -> scalars 0, 2, amg_inner_alpha:
                  rows,cols,values,p,q,t,scalars,status,parent_status,spd */
  KH_INNER_FIRST_STEP,
  KH_INNER_FIRST_UPDATE,
  KH_INNER_BETA_STEP,
  KH_INNER_SECOND_STEP,
  /* KH_AMG_SMOOTH with one wavefront per row: the lanes gather the row and
     form its products and their errors in parallel, one lane sums them in
     entry order (no reference counterpart; same arguments, same bits). For
     levels with long rows. */
  KH_AMG_SMOOTH_ROWWAVE,
  /* KH_FUSED_COARSE_SWEEPS with the operator read once per launch (no
     reference counterpart; same bits): each sweep forms every entry's term
     and product error in parallel (`cap` entries per thread held in
     registers: 8 or 16), then each thread sums its rows in entry order
     (table: first row and end row per thread, built by
     kh_resident_partition). rows,cols,values,b,inverse,x,status,sweeps,
     table,cap (n = rows, at most KH_RESIDENT_MAX_ROWS; at most 16 x
     KH_RESIDENT_THREADS entries). */
  KH_COARSE_SWEEPS_RESIDENT,
  /* KH_CSR_RESIDUAL_GUARDED and KH_FUSED_PROLONG_SMOOTH with one wavefront
     per row, as KH_AMG_SMOOTH_ROWWAVE (no reference counterpart; same
     arguments, same bits), for levels with long rows. */
  KH_CSR_RESIDUAL_ROWWAVE,
  KH_PROLONG_SMOOTH_ROWWAVE,
  KH_KERNEL_COUNT_MARK
} kh_kernel;
#define KH_COARSE_SWEEPS_MAX_ROWS 16384
#define KH_WIDE_ROW 16
#define KH_RESIDENT_THREADS 512
#define KH_RESIDENT_MAX_ROWS 4096
#define KH_RESIDENT_TABLE_INTS (2 * KH_RESIDENT_THREADS)

/* One system of the batched BiCGStab-Jacobi launches: its operator, its
 * vectors (x, r, q, p, z, s, t, zs, shadow; n floats each), its Jacobi
 * inverse, its 24 pair scalars and 3-int status. */
#define KH_BATCH_MAX_LANES 4
typedef struct kh_bicg_lane {
  const int32_t* row_ptr;
  const int32_t* col;
  const float* val;
  float* x;
  float* r;
  float* q;
  float* p;
  float* z;
  float* s;
  float* t;
  float* zs;
  const float* shadow;
  const float* inverse;
  float* scalars;
  int32_t* status;
} kh_bicg_lane;

/* Learned coarse correction: hidden width, gradient partial slots, layers. */
#define KH_LEARNED_HIDDEN 64
#define KH_LEARNED_SLOTS 256
#define KH_LEARNED_MAX_LAYERS 4

/* One level of KH_FUSED_COARSE_TAIL: the level's CSR operator (rows, nnz) and
 * l1 inverse diagonal, and the aggregation transfer to the next level (ignored
 * on the last). Levels run in LDS, so the tail accepts at most
 * KH_TAIL_MAX_LEVELS levels of at most KH_TAIL_MAX_ROWS rows each, with
 * 2 * (sum of rows) at most KH_TAIL_LDS_FLOATS and
 * 2 * (sum of rows) + max(rows) + max(3 * rows + 1 + 2 * nnz) at most
 * KH_TAIL_POOL_WORDS; a table outside these limits makes the kernel a no-op. */
#define KH_TAIL_MAX_LEVELS 16
#define KH_TAIL_MAX_ROWS 8192
#define KH_TAIL_LDS_FLOATS 24576
typedef struct kh_tail_level {
  const int32_t* row_ptr;
  const int32_t* col;
  const float* val;
  const float* inverse;
  const int32_t* member_offsets;
  const int32_t* members;
  const int32_t* aggregate;
  int32_t rows;
  int32_t coarse_rows;
  int32_t nnz;
  int32_t pad;
} kh_tail_level;
/* The tail's LDS pool in 32-bit words: the iterates and right-hand sides of
 * all its levels, one sweep's new values, and the operator of the level being
 * swept (row starts, columns, values, l1 inverse, aggregate map), copied in on
 * each visit. */
#define KH_TAIL_POOL_WORDS 40448

/* Device::launch (device.hpp:205; cuda_device.cu:132-140): runs kernel `kernel`
 * for global indices [0, n). n == 0 is a no-op; n > INT32_MAX, an unknown id or
 * arguments of the wrong number or kind are errors. */
int kh_launch(kh_device* dev, uint32_t kernel, size_t n, const kh_arg* args, size_t nargs);
/* Name of a kernel id, or NULL. */
const char* kh_kernel_name(uint32_t kernel);

/* Host reductions: Device::sum/dot/max_abs and host_reductions
 * (device.hpp:157-184, 227-233). FP64 accumulation in a fixed order; max is
 * NaN-propagating. All requests share one host wait. */
enum { KH_REDUCE_SUM = 0, KH_REDUCE_DOT = 1, KH_REDUCE_MAX_ABS = 2 };
typedef struct kh_reduction {
  uint32_t kind;
  uint32_t pad;
  const float* x;
  const float* y;
  size_t n;
} kh_reduction;
int kh_reduce_host(kh_device* dev, const kh_reduction* requests, size_t count, double* out);

/* Device::reduce_into (device.hpp:186-202, 211-225): asynchronous FP64 sum,
 * dot or fused dot pair, stored as normalized (hi, lo) float pairs at
 * dst[offset..offset+1] (and +2..3 for KH_INTO_DOT2). dst must not alias inputs. */
enum { KH_INTO_SUM = 0, KH_INTO_DOT = 1, KH_INTO_DOT2 = 2 };
int kh_reduce_into(kh_device* dev, int mode, const float* x, const float* y, const float* u,
                   const float* v, size_t n, float* dst, size_t offset);

/* kh_reduce_into without a device handle and without allocating (no
 * reference counterpart). For callers that cannot allocate device memory:
 * code under HIP graph capture, serving loops, runtimes that own their own
 * streams. Same arithmetic, order and output as kh_reduce_into; the
 * per-workgroup partials live in a fixed 32 KiB array that each GPU holds for
 * the life of the process, so the call makes no hipMalloc/hipFree. `stream` is
 * a hipStream_t of the current device (NULL: its null stream); the work is
 * asynchronous on it. n above KH_STATIC_REDUCE_MAX_N is an error. All calls
 * on one GPU share the partials array: calls that can overlap on that GPU
 * (different streams without ordering between them) race. */
#define KH_STATIC_REDUCE_MAX_N 2147483647u
int kh_reduce_into_static(void* stream, int mode, const float* x, const float* y, const float* u,
                          const float* v, size_t n, float* dst, size_t offset);

/* Graph capture of a fixed launch sequence (no reference counterpart: the
 * reference replays nothing on CUDA). Between begin and end, entry points that
 * enqueue work are recorded instead of executed; blocking calls are errors. */
typedef struct kh_graph kh_graph;
int kh_graph_begin(kh_device* dev);
int kh_graph_end(kh_device* dev, kh_graph** out);
int kh_graph_launch(kh_device* dev, kh_graph* graph);
void kh_graph_destroy(kh_graph* graph);

/* Per-kernel timing with HIP events on the device stream (no reference
 * counterpart): while enabled, every launch of `kernel` is bracketed by events.
 * kh_profile_read synchronizes and returns the launch count and summed ms. */
int kh_profile_enable(kh_device* dev, uint32_t kernel, int enable);
/* Restricts timing of `kernel` to launches over exactly n indices (0: all). */
int kh_profile_filter(kh_device* dev, uint32_t kernel, size_t n);
int kh_profile_read(kh_device* dev, uint32_t kernel, long* launches, double* total_ms);
int kh_profile_reset(kh_device* dev);
/* The reference's logical-bytes model of device traffic (no reference
 * counterpart as an entry point; bench/bench_fv.cpp:68-93 counts the same way):
 * while enabled, each kernel launch adds the full size of the allocation behind
 * every buffer argument, reduce_into adds 4n / 8n / 16n bytes (sum / dot /
 * dot2) and a host reduction 8n (dot) or 4n. Returns the totals since the last
 * call, resets them and sets counting on or off. Launches recorded into a
 * graph are not counted. */
int kh_logical_bytes(kh_device* dev, int enable, double* bytes, long* launches);
/* Times `repeats` back-to-back launches of one kernel (no reference
 * counterpart): one HIP event before the first and one after the last on the
 * device stream, then a synchronization; *total_ms is the elapsed time of the
 * whole batch (kernel execution plus the gaps between launches). */
int kh_time_launches(kh_device* dev, uint32_t kernel, size_t n, const kh_arg* args, size_t nargs,
                     int repeats, double* total_ms);
/* The row partition KH_COARSE_SWEEPS_RESIDENT reads (no reference
 * counterpart; host only, no device call): from a CSR pattern of `rows` rows
 * (row_ptr[0] = 0, every row non-empty, rows <= KH_RESIDENT_MAX_ROWS, at most
 * 16 x KH_RESIDENT_THREADS entries; cols is not read), fills table
 * (KH_RESIDENT_TABLE_INTS ints) and *cap (8 or 16). Returns 0, or 1 when the
 * level does not fit (then nothing is written). */
int kh_resident_partition(const int32_t* row_ptr, const int32_t* cols, int32_t rows, int32_t* table,
                          int32_t* cap);

#ifdef __cplusplus
}
#endif

#endif /* KUTTA_HIP_H */
