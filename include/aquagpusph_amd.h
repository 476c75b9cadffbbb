/*
 * aquagpusph_amd — MI355X-native (gfx950, HIP) implementation of the
 * AQUAgpusph per-timestep SPH pipeline.
 *
 * Plain C ABI: device pointers, sizes and a HIP stream passed as void*.
 * Every entry point returns 0 on success or a negative AQ_E* code; the text
 * of the last error on the calling thread is aq_last_error().  Nothing here
 * throws across the boundary.
 *
 * Conventions (mirroring the reference's variable contract,
 * aquagpusph/Variable.cpp:1548-1575 with addr_bits = 32):
 *   - `vec` arrays hold float2 (2D) or float4 (3D, w carried) per particle;
 *   - indices (`usize`, `size_t*` variables) are uint32;
 *   - `imove` is int32, `iset` is uint32;
 *   - per-set arrays (refd, visc_dyn, delta) are float[n_sets];
 *   - scalars that can change during a run (dt, t, n_cells, r_min, cs, g,
 *     ...) are passed as DEVICE pointers so that a step can be replayed from
 *     a captured HIP graph without re-binding;
 *   - preset definitions (H, CONW, CONF, DIMS, Laplacian formulation) are
 *     compile-time constants in the reference and are passed by value in
 *     aq_defs.
 * Each particle kernel launches over N particles and reproduces the
 * reference kernel's early exits on imove.
 */
#ifndef AQUAGPUSPH_AMD_H
#define AQUAGPUSPH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AQ_API __attribute__((visibility("default")))

/* error codes */
#define AQ_OK 0
#define AQ_EINVAL -1    /* invalid argument */
#define AQ_EHIP -2      /* HIP runtime error */
#define AQ_ENOMEM -3    /* allocation failure or capacity exceeded */
#define AQ_EPARSE -4    /* XML / expression parse error */
#define AQ_ENOTFOUND -5 /* unknown variable, tool or script */
#define AQ_EASSERT -6   /* an assert tool failed on the device */
#define AQ_EUNSUP -7    /* feature not supported by this build */

/* Laplacian formulation (cfd.xml __LAP_FORMULATION__) */
#define AQ_LAP_MONAGHAN 1
#define AQ_LAP_MORRIS 2

/* Preset definitions, as the reference passes them to its kernels with -D
 * after "%#G" formatting (aquagpusph/CalcServer/CalcServer.cpp:241-260). */
typedef struct aq_defs {
    float H;    /* kernel length h */
    float CONW; /* 1/h^dims */
    float CONF; /* 1/h^(dims+2) */
    float DIMS; /* dims as a float */
    int lap_formulation; /* AQ_LAP_MONAGHAN or AQ_LAP_MORRIS */
} aq_defs;

/* ---- library ---------------------------------------------------------- */
AQ_API const char* aq_version(void);
AQ_API const char* aq_strerror(int code);
AQ_API const char* aq_last_error(void);
/* Select the HIP device used by subsequent calls on this thread. */
AQ_API int aq_set_device(int device);

/* ---- link-list: aquagpusph/CalcServer/LinkList.{cpp,cl.in} ------------ */

/* r_min/r_max reductions + LinkList::nCells (LinkList.cpp:185-232,
 * 326-391): component-wise min/max of r[0..N) into r_min/r_max (device vec
 * scalars), then n_cells.xyz = (uint)((max-min)/cell_length) + 6 and
 * n_cells.w = x*y*z into the device uint4 n_cells.  If n_cells.w exceeds
 * ihoc_capacity, *status (device uint) gets bit 0 set; the following
 * link-list kernels then do nothing.  workspace: aq_linklist_workspace(). */
AQ_API int aq_linklist_grid(int dims, const float* r, uint32_t N,
                            float* r_min, float* r_max,
                            const float* support, const float* h,
                            uint32_t* n_cells, uint32_t ihoc_capacity,
                            uint32_t* status, void* workspace, void* stream);

/* iCell kernel, LinkList.cl.in:54-85. */
AQ_API int aq_linklist_icell(int dims, uint32_t* icell, const float* r,
                             uint32_t N, const float* r_min,
                             const float* support, const float* h,
                             const uint32_t* n_cells, const uint32_t* status,
                             void* stream);

/* iHoc + linkList kernels, LinkList.cl.in:32-42 and 92-113, on sorted icell.
 * Launches over ihoc_capacity cells; only cells < n_cells.w are written. */
AQ_API int aq_linklist_ihoc(const uint32_t* icell, uint32_t* ihoc,
                            uint32_t N, const uint32_t* n_cells,
                            uint32_t ihoc_capacity, const uint32_t* status,
                            void* stream);

/* Whole LinkList tool (LinkList::_execute, LinkList.cpp:326-494) on the
 * positions r: grid, iCell, stable radix sort of icell (id_unsorted = perm,
 * id_sorted = inverse perm), iHoc, linkList. */
AQ_API int aq_linklist(int dims, const float* r, uint32_t N, float* r_min,
                       float* r_max, const float* support, const float* h,
                       uint32_t* n_cells, uint32_t* icell,
                       uint32_t* id_unsorted, uint32_t* id_sorted,
                       uint32_t* ihoc, uint32_t ihoc_capacity,
                       uint32_t* status, void* workspace, void* stream);
/* Bytes of device workspace aq_linklist / aq_linklist_grid need for N. */
AQ_API size_t aq_linklist_workspace(uint32_t N);

/* aq_linklist without a workspace argument and without any allocation.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The workspace lives in a fixed device array of the
 * library (9 MiB per GPU, held until the process ends), so N is limited to
 * AQ_LINKLIST_NOALLOC_MAX_N particles (larger N: AQ_EINVAL, nothing
 * launched).  Every call on one GPU shares that array: calls that overlap
 * on one GPU (other streams, other threads) race and must be serialised by
 * the caller.  Outputs and their meaning are those of aq_linklist. */
#define AQ_LINKLIST_NOALLOC_MAX_N 1048576u
AQ_API int aq_linklist_noalloc(int dims, const float* r, uint32_t N,
                               float* r_min, float* r_max,
                               const float* support, const float* h,
                               uint32_t* n_cells, uint32_t* icell,
                               uint32_t* id_unsorted, uint32_t* id_sorted,
                               uint32_t* ihoc, uint32_t ihoc_capacity,
                               uint32_t* status, void* stream);

/* ---- radix sort: aquagpusph/CalcServer/RadixSort.{cpp,cl.in} ---------- */

/* Stable LSD radix sort of n uint32 keys on their low key_bits bits
 * (RadixSort::_execute, RadixSort.cpp:129-303).  keys is sorted in place;
 * perm[k] = original index of sorted key k (id_unsorted); inv_perm[perm[k]]
 * = k (id_sorted, RadixSort.cl.in:313-323).  inv_perm may be NULL.
 * workspace: aq_radix_sort_workspace(n). */
AQ_API int aq_radix_sort(uint32_t* keys, uint32_t* perm, uint32_t* inv_perm,
                         uint32_t n, unsigned key_bits, void* workspace,
                         void* stream);
AQ_API size_t aq_radix_sort_workspace(uint32_t n);
/* Key bits the reference processes for max_val (RadixSort.cpp:152-176). */
AQ_API unsigned aq_radix_key_bits(uint64_t max_val);

/* ---- reorder: resources/Scripts/basic/Sort.cl:57-124 ------------------ */

/* dst[k] = src[perm[k]] for k < N, element size esize bytes (1..16, power of
 * two).  With perm = id_unsorted this equals the reference scatter
 * x[id_sorted[i]] = x_in[i]. */
AQ_API int aq_permute(const void* src, void* dst, size_t esize,
                      const uint32_t* perm, uint32_t N, void* stream);

/* aquagpusph/CalcServer/UnSort.cl.in:31-42: out[id[i]] = in[i] for i < N
 * (back to the original particle order), elem_bytes a multiple of 4
 * (4, 8, 16 or 64). */
AQ_API int aq_unsort(const uint32_t* id, const void* in, void* out,
                     uint32_t N, uint32_t elem_bytes, void* stream);

/* ---- streaming particle kernels --------------------------------------- */

/* improved Euler predictor, basic/time_scheme/improved_euler.cl:75-103 */
AQ_API int aq_predictor(int dims, const int32_t* imove, const float* r,
                        const float* u, const float* dudt, const float* rho,
                        const float* drhodt, float* r_in, float* u_in,
                        float* dudt_in, float* rho_in, float* drhodt_in,
                        uint32_t N, const float* dt, void* stream);

/* improved Euler corrector, basic/time_scheme/improved_euler.cl:125-147 */
AQ_API int aq_corrector(int dims, const int32_t* imove, float* r, float* u,
                        const float* dudt, float* rho, const float* drhodt,
                        const float* dudt_in, const float* drhodt_in,
                        uint32_t N, const float* dt, void* stream);

/* basic/Domain.cl:48-90 */
AQ_API int aq_domain(int dims, int32_t* imove, float* r_in, float* u_in,
                     float* dudt_in, float* m, uint32_t N,
                     const float* domain_min, const float* domain_max,
                     void* stream);

/* basic/EOS.cl:57-73 */
AQ_API int aq_eos(const uint32_t* iset, const int32_t* imove,
                  const float* rho, float* p, const float* refd, uint32_t N,
                  const float* cs, const float* p0, void* stream);

/* basic/Binormal.cl:37-53 */
AQ_API int aq_binormal(int dims, const float* normal, float* tangent,
                       float* binormal, uint32_t N, void* stream);

/* cfd/Rates.cl:55-77 */
AQ_API int aq_rates(int dims, const uint32_t* iset, const int32_t* imove,
                    const float* rho, const float* grad_p, const float* lap_u,
                    const float* div_u, float* dudt, float* drhodt,
                    const float* visc_dyn, uint32_t N, const float* g,
                    void* stream);

/* cfd/Energy/EnergyKin.cl:40-56: energy_kin = m u.u / 2 (fluid; 0 else) */
AQ_API int aq_energy_kin(int dims, float* energy_kin, const int32_t* imove,
                         const float* u, const float* m, uint32_t N,
                         void* stream);

/* cfd/Energy/Energy.cl:59-86 (power): kinetic, potential and
 * compressibility energy rates per fluid particle; g on the device */
AQ_API int aq_energy_power(int dims, float* energy_dekdt, float* energy_depdt,
                           float* energy_decdt, const int32_t* imove,
                           const float* u, const float* rho, const float* m,
                           const float* p, const float* dudt,
                           const float* drhodt, uint32_t N, const float* g,
                           void* stream);

/* cfd/Energy/Energy.cl:113-142 (energy): kinetic, potential and
 * compressibility energies per fluid particle; g and cs on the device */
AQ_API int aq_energy(int dims, float* energy_ek, float* energy_ep,
                     float* energy_ec, const uint32_t* iset,
                     const int32_t* imove, const float* r, const float* u,
                     const float* rho, const float* m, const float* refd,
                     uint32_t N, const float* g, const float* cs,
                     void* stream);

/* basic/deltaSPH.cl:330-350 (cfd/deltaSPH.cl: fluid only) */
AQ_API int aq_delta_rate(const uint32_t* iset, const int32_t* imove,
                         const float* rho, const float* lap_p, float* drhodt,
                         const float* refd, const float* delta, uint32_t N,
                         const float* dt, void* stream);

/* cfd/TimeStep.cl:56-77 */
AQ_API int aq_timestep(int dims, const int32_t* imove, const float* u,
                       float* dt_var, uint32_t N, const float* dt,
                       const float* dt_min, const float* courant,
                       const float* dt_Ma, const float* h, void* stream);

/* cfd/SensorsRenormalization.cl:42-67 */
AQ_API int aq_sensors_renorm(int dims, const int32_t* imove,
                             const float* shepard, float* u, float* rho,
                             float* p, uint32_t N, void* stream);

/* cfd/Boundary/BI/Shepard.cl:173-195 (apply) */
AQ_API int aq_bi_shepard_apply(int dims, const int32_t* imove,
                               const float* shepard, float* grad_p,
                               float* lap_u, float* div_u, uint32_t N,
                               void* stream);

/* cfd/Boundary/BI/GradP.cl:55-83 (freeslip) */
AQ_API int aq_bi_gradp(int dims, const uint32_t* iset, const int32_t* imove,
                       const float* shepard, const float* rho,
                       const float* lap_u, const float* dudt, float* grad_p,
                       const float* visc_dyn, const float* refd, uint32_t N,
                       const float* g, void* stream);

/* cfd/Boundary/BI/InterpolationShepard.cl:48-76 */
AQ_API int aq_bi_interp_shepard(const uint32_t* iset, const int32_t* imove,
                                const float* shepard, float* rho, float* p,
                                const float* refd, uint32_t N,
                                const float* cs, const float* p0,
                                void* stream);

/* ---- pair (neighbour-walk) kernels ------------------------------------
 * All take the link-list (icell, ihoc, n_cells) of the sorted particles and
 * visit neighbours in the reference order (resources/Scripts/types/3D.h:
 * 197-219, 2D.h:174-193), so FP32 sums associate exactly as there. */

/* basic/Shepard.cl:76-125.  cfd_variant != 0 selects cfd/Shepard.cl
 * (EXCLUDED_PARTICLE imove != 1), else the basic default (imove >= 3). */
AQ_API int aq_shepard(int dims, const int32_t* imove, const float* r,
                      const float* rho, const float* m, float* shepard,
                      uint32_t N, const uint32_t* icell, const uint32_t* ihoc,
                      const uint32_t* n_cells, aq_defs defs, int cfd_variant,
                      void* stream);

/* cfd/Interactions.cl:60-145 */
AQ_API int aq_interactions(int dims, const int32_t* imove, const float* r,
                           const float* u, const float* rho, const float* m,
                           const float* p, float* grad_p, float* lap_u,
                           float* div_u, uint32_t N, const uint32_t* icell,
                           const uint32_t* ihoc, const uint32_t* n_cells,
                           aq_defs defs, void* stream);

/* basic/deltaSPH.cl:191-242 (lapp), cfd variant (fluid only) */
AQ_API int aq_lapp(int dims, const int32_t* imove, const float* r,
                   const float* rho, const float* m, const float* p,
                   float* lap_p, uint32_t N, const uint32_t* icell,
                   const uint32_t* ihoc, const uint32_t* n_cells,
                   aq_defs defs, void* stream);

/* cfd/Sensors.cl:57-130 */
AQ_API int aq_sensors(int dims, const int32_t* imove, const float* r,
                      const float* m, float* u, float* rho, float* p,
                      uint32_t N, const uint32_t* icell, const uint32_t* ihoc,
                      const uint32_t* n_cells, aq_defs defs, void* stream);

/* cfd/Boundary/BI/Shepard.cl:61-150 (compute) */
AQ_API int aq_bi_shepard_compute(int dims, const int32_t* imove,
                                 const float* r, const float* normal,
                                 const float* tangent, const float* binormal,
                                 const float* m, float* shepard, uint32_t N,
                                 const uint32_t* icell, const uint32_t* ihoc,
                                 const uint32_t* n_cells, aq_defs defs,
                                 void* stream);

/* cfd/Boundary/BI/LapU.cl:58-123 (freeslip) */
AQ_API int aq_bi_lapu(int dims, const int32_t* imove, const float* r,
                      const float* u, const float* rho, const float* m,
                      float* lap_u, uint32_t N, const uint32_t* icell,
                      const uint32_t* ihoc, const uint32_t* n_cells,
                      aq_defs defs, void* stream);

/* cfd/Boundary/BI/Interpolation.cl:54-107 */
AQ_API int aq_bi_interp(int dims, const int32_t* imove, const float* r,
                        const float* m, const float* rho,
                        const float* grad_p, float* p, uint32_t N,
                        const uint32_t* icell, const uint32_t* ihoc,
                        const uint32_t* n_cells, aq_defs defs, void* stream);

/* cfd/Boundary/BI/Interactions.cl:51-133 */
AQ_API int aq_bi_interactions(int dims, const int32_t* imove, const float* r,
                              const float* normal, const float* u,
                              const float* rho, const float* m,
                              const float* p, float* grad_p, float* div_u,
                              uint32_t N, const uint32_t* icell,
                              const uint32_t* ihoc, const uint32_t* n_cells,
                              aq_defs defs, void* stream);

/* basic/neighs.cl:52-91 */
AQ_API int aq_neighs(int dims, const int32_t* imove, uint32_t* n_neighs,
                     const uint32_t* neighs_limit, uint32_t N,
                     const uint32_t* icell, const uint32_t* ihoc,
                     const uint32_t* n_cells, void* stream);

/* ---- reductions: aquagpusph/CalcServer/Reduction.{cpp,cl.in} ---------- */
#define AQ_RED_MIN 0
#define AQ_RED_MAX 1
#define AQ_RED_SUM 2
#define AQ_T_FLOAT 0 /* float */
#define AQ_T_INT 1   /* int32 */
#define AQ_T_UINT 2  /* uint32 */
#define AQ_T_DOUBLE 3    /* double */
#define AQ_T_FLOAT_F64 4 /* float elements reduced in double, double out */
/* Reduce n elements of `ncomp` components each (1 for scalars, 2/4 for
 * vec) component-wise with op into out[ncomp] (device).  identity is the
 * reference's "null" value per component (host array of ncomp values, as
 * float/int32/uint32/double bits per the output type of `type`).
 * workspace: aq_reduce_workspace(). */
AQ_API int aq_reduce(int op, int type, int ncomp, const void* in, uint32_t n,
                     void* out, const void* identity, void* workspace,
                     void* stream);
AQ_API size_t aq_reduce_workspace(uint32_t n, int ncomp);

/* ---- simulation runtime: CalcServer + XML problem setup --------------
 * The counterpart of the reference executable's CalcServer
 * (aquagpusph/CalcServer/CalcServer.cpp:139-621) behind the same XML case
 * and preset files (aquagpusph/InputOutput/State.cpp): variables, preset
 * definitions, the ordered tool list and the particle sets.  Kernel tools
 * resolve their `path` + `entry_point` to the HIP kernels above. */
typedef struct aq_sim aq_sim;

/* Parse the case (dims 2 or 3, like the reference's -d option), register
 * the variables, build and set up every tool and load the particle files. */
AQ_API int aq_sim_create(const char* xml_path, int dims, aq_sim** out);
AQ_API void aq_sim_destroy(aq_sim* sim);
/* Run nsteps time steps by walking the tool list (CalcServer::update). */
AQ_API int aq_sim_step(aq_sim* sim, int nsteps);
/* Same, replaying one captured HIP graph per step after a first eager step. */
AQ_API int aq_sim_step_graph(aq_sim* sim, int nsteps);
AQ_API int aq_sim_sync(aq_sim* sim);
AQ_API void* aq_sim_stream(aq_sim* sim);
AQ_API uint32_t aq_sim_n_particles(aq_sim* sim);
/* Type (canonical reference name), length and element size of a variable. */
AQ_API int aq_sim_var_info(aq_sim* sim, const char* name, char* type_buf,
                           size_t type_len, size_t* length,
                           size_t* elem_bytes);
/* Copy a variable (array or scalar) to/from host memory. */
AQ_API int aq_sim_get(aq_sim* sim, const char* name, void* host,
                      size_t bytes);
AQ_API int aq_sim_set(aq_sim* sim, const char* name, const void* host,
                      size_t bytes);
/* Device address of an array variable or of a scalar's slot. */
AQ_API void* aq_sim_device_ptr(aq_sim* sim, const char* name);
AQ_API int aq_sim_n_tools(aq_sim* sim);
AQ_API const char* aq_sim_tool_name(aq_sim* sim, int i);
AQ_API const char* aq_sim_tool_type(aq_sim* sim, int i);
/* Formatted value of a preset definition (as the kernels receive it). */
AQ_API int aq_sim_definition(aq_sim* sim, const char* name, char* buf,
                             size_t len);
/* Mean GPU milliseconds per tool over nsteps eager steps, into ms[n_tools]
 * (the report_performance counterpart, CalcServer/Reports/Performance.cpp). */
AQ_API int aq_sim_profile_tools(aq_sim* sim, int nsteps, double* ms);
/* ---- LES eddy-viscosity network (no counterpart in the reference) ----
 * A 7 -> 64 -> 64 -> 1 perceptron with GELU hidden units; BF16 matrix
 * instructions, FP32 accumulation.  params: aq_mlp_nparam() floats, FP32
 * master copy laid out W1[64][16] (inputs padded from 7; columns 7..15
 * zero), b1[64], W2[64][64] (out, in), b2[64], w3[64], b3.  x: m rows of
 * 8 floats (7 inputs + 1 ignored). */
AQ_API int aq_mlp_nparam(void);
AQ_API size_t aq_mlp_workspace(uint32_t m);
/* y[i] = network(x[i]) for i < m (device pointers). */
AQ_API int aq_mlp_forward(const float* x, uint32_t m, const float* params,
                          float* y, void* stream);
/* grad[nparam + 1]: the gradient of the mean squared error over the m
 * samples and, last, the loss itself.  Deterministic (fixed summation
 * order).  workspace: aq_mlp_workspace(m) bytes. */
AQ_API int aq_mlp_grad(const float* x, const float* target, uint32_t m,
                       const float* params, float* grad, void* workspace,
                       void* stream);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8; step counts from 1) on
 * the master copy from grad (its first nparam entries). */
AQ_API int aq_mlp_adam(float* params, const float* grad, float* adam_m,
                       float* adam_v, int step, float lr, void* stream);

/* Autotuner (no counterpart in the reference): times the launch
 * configurations of the fused cell walk (segment width 8/16, one or eight
 * cell hand-out ranges, resident or one-per-CU grid) over nsteps eager steps
 * each on this case, restores the state after each, keeps the fastest and
 * drops captured graphs.  A text report (one line per configuration and the
 * choice) goes to report[len]. */
AQ_API int aq_sim_autotune(aq_sim* sim, int nsteps, char* report,
                           size_t len);
/* The AQUAgpusph main loop (aquagpusph/main.cpp:160-180, TimeManager.cpp:
 * 103-150): steps (graph replays when graph != 0) until the Timing end
 * criteria, writing each ParticlesSet's <Save> files, in the original
 * particle order, and the state file AQUAgpusph.save.N.xml (in the working
 * directory) at every output point; at most max_steps steps when
 * max_steps > 0.  *steps receives the number of steps run. */
AQ_API int aq_sim_run(aq_sim* sim, long max_steps, int graph, long* steps);
/* Write the <Save> files and the state file now (FileManager.cpp:146-155). */
AQ_API int aq_sim_save(aq_sim* sim);
/* Mean GPU milliseconds of the fused fluid neighbour walk kernel alone
 * (HIP events around its launch) over the last aq_sim_profile_tools call;
 * -1 when the pipeline has no fused walk.  No counterpart in the reference:
 * the fused walk replaces three of its kernel tools. */
AQ_API int aq_sim_walk_ms(aq_sim* sim, double* ms);
/* Parse a case without using the GPU and describe the resolved pipeline:
 * "name<TAB>type<TAB>key" per tool in execution order, then "#var", "#def"
 * and "#N" lines.  *needed receives the buffer size required. */
AQ_API int aq_xml_resolve(const char* xml_path, int dims, char* out,
                          size_t len, size_t* needed);
/* Hardware check of the fused walk's division and square root sequences
 * without range scaling (no counterpart in the reference, which leaves
 * division to the OpenCL compiler): n random operand pairs over the whole
 * window where the walk uses them are divided (and square-rooted) both ways
 * on the current device; mismatches[0] / [1] receive the number of
 * quotients / square roots whose bit patterns differ from IEEE a / b and
 * sqrtf (0 expected).  Synchronous. */
AQ_API int aq_check_nr_arith(uint64_t n, uint64_t seed, uint64_t* mismatches);
/* The float atan (op 0) and acospi (op 1) of the boundary-integral kernel
 * terms, y[k] = f(x[k]) for device arrays of n floats, on `stream` (the
 * reference takes atan / acospi from its OpenCL runtime:
 * KernelFunctions/Wendland2D.hcl, Wendland3D.hcl:137-185).  For checking
 * the device sequence against the CPU one. */
AQ_API int aq_clmath_eval(int op, const float* x, float* y, uint32_t n,
                          void* stream);
/* Registry keys ("cfd/Interactions.cl:entry", ...) of the HIP kernels. */
AQ_API int aq_sim_n_registered_kernels(void);
AQ_API const char* aq_sim_registered_kernel(int i);

#ifdef __cplusplus
}
#endif

#endif /* AQUAGPUSPH_AMD_H */
