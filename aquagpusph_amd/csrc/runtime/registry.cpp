// Kernel-tool registry: (script path under Scripts/, entry point) -> HIP
// launcher with the reference kernel's arguments bound by variable name
// (aquagpusph/CalcServer/Kernel.cpp:464-556 binds OpenCL kernel arguments by
// name the same way).  A preset keeps its `type="kernel" path=... entry_point=`
// lines; only the implementation behind the key changes.
#include <algorithm>
#include <filesystem>
#include <map>

#include "sim.hpp"

namespace aq {

void permute(const void* src, void* dst, size_t esize, const uint32_t* perm,
             uint32_t n, hipStream_t s);
void permute_many(const void* const* src, void* const* dst,
                  const size_t* esize, int nf, const uint32_t* perm,
                  uint32_t n, hipStream_t s);
void launch_predictor(int, const int32_t*, const float*, const float*,
                      const float*, const float*, const float*, float*, float*,
                      float*, float*, float*, uint32_t, const float*,
                      const uint32_t*, hipStream_t);
void launch_corrector(int, const int32_t*, float*, float*, const float*,
                      float*, const float*, const float*, const float*,
                      uint32_t, const float*, const uint32_t*, hipStream_t);
void launch_domain(int, int32_t*, float*, float*, float*, float*, uint32_t,
                   const float*, const float*, hipStream_t);
void launch_eos(const uint32_t*, const int32_t*, const float*, float*,
                const float*, uint32_t, const float*, const float*,
                hipStream_t);
void launch_binormal(int, const float*, float*, float*, uint32_t, hipStream_t);
void launch_tait_eos(const uint32_t*, const int32_t*, const float*, float*,
                     const float*, uint32_t, const float*, const float*,
                     const float*, hipStream_t);
void launch_bi_interp_shepard_tait(const uint32_t*, const int32_t*,
                                   const float*, float*, float*, const float*,
                                   uint32_t, const float*, const float*,
                                   const float*, hipStream_t);
void launch_rates(int, const uint32_t*, const int32_t*, const float*,
                  const float*, const float*, float*, float*, const float*,
                  uint32_t, const float*, hipStream_t);
void launch_delta_rate(const uint32_t*, const int32_t*, const float*,
                       const float*, float*, const float*, const float*,
                       uint32_t, const float*, hipStream_t);
void launch_timestep(int, const int32_t*, const float*, float*, uint32_t,
                     const float*, const float*, const float*, const float*,
                     const float*, hipStream_t);
void launch_sensors_renorm(int, const int32_t*, const float*, float*, float*,
                           float*, uint32_t, hipStream_t);
void launch_portal_mirror(int, float*, int32_t*, uint32_t*, uint32_t,
                          const float*, const float*, const float*,
                          const float*, const uint32_t*, aq_defs, hipStream_t);
void launch_portal_unmirror(int, float*, const int32_t*, uint32_t,
                            const float*, const float*, hipStream_t);
void launch_portal_teleport(int, float*, uint32_t, const float*, const float*,
                            const float*, hipStream_t);
void launch_portal_shepard(int, const int32_t*, const int32_t*, const float*,
                           const float*, const float*, float*, uint32_t,
                           const uint32_t*, const uint32_t*, const uint32_t*,
                           const uint32_t*, aq_defs, float, hipStream_t);
void launch_portal_interactions(int, const int32_t*, const int32_t*,
                                const float*, const float*, const float*,
                                const float*, const float*, float*, float*,
                                float*, uint32_t, const uint32_t*,
                                const uint32_t*, const uint32_t*,
                                const uint32_t*, aq_defs, float, hipStream_t);
void launch_portal_bi(int, const int32_t*, const int32_t*, const float*,
                      const float*, const float*, const float*, const float*,
                      const float*, float*, float*, uint32_t, const uint32_t*,
                      const uint32_t*, const uint32_t*, const uint32_t*,
                      aq_defs, float, hipStream_t);
void launch_portal_lapp(int, const int32_t*, const int32_t*, const float*,
                        const float*, const float*, const float*, float*,
                        float*, uint32_t, const uint32_t*, const uint32_t*,
                        const uint32_t*, const uint32_t*, aq_defs, float,
                        hipStream_t);
void launch_portal_lapp_corr(int, const int32_t*, const int32_t*,
                             const float*, const float*, const float*,
                             const float*, float*, uint32_t, const uint32_t*,
                             const uint32_t*, const uint32_t*,
                             const uint32_t*, aq_defs, float, hipStream_t);
void launch_portal_mls(int, const int32_t*, const int32_t*, const float*,
                       const float*, const float*, float*, const uint32_t*,
                       uint32_t, const uint32_t*, const uint32_t*,
                       const uint32_t*, const uint32_t*, aq_defs, float,
                       hipStream_t);
void launch_predictor64(int, const int32_t*, const float*, const double*,
                        const float*, const float*, const float*,
                        const float*, float*, double*, float*, float*, float*,
                        float*, uint32_t, const float*, const uint32_t*,
                        hipStream_t);
void launch_corrector64(int, const int32_t*, float*, double*, float*,
                        const float*, float*, const float*, const float*,
                        const float*, uint32_t, const float*, const uint32_t*,
                        hipStream_t);
void launch_fluid_mass(const int32_t*, const float*, float*, uint32_t,
                       hipStream_t);
void launch_bt_eos(const uint32_t*, const int32_t*, const float*, float*,
                   const float*, const float*, uint32_t, const float*,
                   hipStream_t);
void launch_bt_rates(int, const uint32_t*, const int32_t*, const float*,
                     const float*, const float*, float*, float*, const float*,
                     const float*, const float*, const float*, const float*,
                     uint32_t, hipStream_t);
void launch_bt_timestep(int, const uint32_t*, const int32_t*, const float*,
                        float*, uint32_t, const float*, const float*,
                        const float*, const float*, const float*,
                        hipStream_t);
void launch_bt_dt_min(const uint32_t*, const float*, uint32_t, uint32_t,
                      float*, hipStream_t);
void launch_bt_predictor(int, const uint32_t*, const int32_t*, const float*,
                         const float*, const float*, const float*,
                         const float*, float*, float*, float*, float*, float*,
                         uint32_t, const float*, hipStream_t);
void launch_bt_corrector(int, const uint32_t*, const int32_t*, float*, float*,
                         const float*, float*, const float*, const float*,
                         const float*, uint32_t, const float*, hipStream_t);
void launch_bt_delta(const uint32_t*, const int32_t*, const float*,
                     const float*, float*, const float*, const float*,
                     uint32_t, const float*, hipStream_t);
void launch_bt_bi_gradp(int, const uint32_t*, const int32_t*, const float*,
                        const float*, const float*, float*, const float*,
                        const float*, const float*, const float*,
                        const float*, const float*, uint32_t, hipStream_t);
void launch_bt_bi_interp_shepard(const uint32_t*, const int32_t*,
                                 const float*, float*, float*, const float*,
                                 const float*, uint32_t, const float*,
                                 hipStream_t);
void launch_bt_time(float*, const float*, uint32_t, hipStream_t);
void launch_les_gradu(int, const int32_t*, const float*, const float*,
                      const float*, const float*, float*, uint32_t,
                      const uint32_t*, const uint32_t*, const uint32_t*,
                      aq_defs, float, hipStream_t);
void launch_les_features(int, const int32_t*, const float*, float*, uint32_t,
                         float, hipStream_t);
void launch_les_nut(int, const int32_t*, const float*, const float*, float*,
                    uint32_t, float, int, float, float, hipStream_t);
void launch_les_interactions(int, const int32_t*, const float*, const float*,
                             const float*, const float*, const float*, float*,
                             uint32_t, const uint32_t*, const uint32_t*,
                             const uint32_t*, aq_defs, float, hipStream_t);
void launch_mlp_forward(const float*, uint32_t, const float*, float*,
                        hipStream_t);
void launch_mp_predictor(int, const float*, const float*, const float*,
                         const float*, const float*, float*, float*, float*,
                         float*, float*, uint32_t, hipStream_t);
void launch_mp_advance(int, const int32_t*, const float*, float*,
                       const float*, const float*, float*, const float*,
                       uint32_t, const float*, hipStream_t);
void launch_mp_advance_r(int, const int32_t*, const float*, float*,
                         const float*, uint32_t, const float*, hipStream_t);
void launch_mp_relax(int, const int32_t*, const float*, float*, const float*,
                     float*, uint32_t, const float*, hipStream_t);
void launch_mp_residuals(int, const int32_t*, const float*, const float*,
                         const float*, const float*, const float*,
                         const float*, const float*, const float*, float*,
                         uint32_t, hipStream_t);
void launch_mp_corrector(int, const int32_t*, const float*, float*,
                         const float*, float*, const float*, const float*,
                         float*, const float*, uint32_t, const float*,
                         hipStream_t);
void launch_energy_kin(int, float*, const int32_t*, const float*,
                       const float*, uint32_t, hipStream_t);
void launch_energy_power(int, float*, float*, float*, const int32_t*,
                         const float*, const float*, const float*,
                         const float*, const float*, const float*, uint32_t,
                         const float*, hipStream_t);
void launch_energy_parts(int, float*, float*, float*, const uint32_t*,
                         const int32_t*, const float*, const float*,
                         const float*, const float*, const float*, uint32_t,
                         const float*, const float*, hipStream_t);
void launch_bi_shepard_apply(int, const int32_t*, const float*, float*,
                             float*, float*, uint32_t, hipStream_t);
void launch_bi_gradp(int, const uint32_t*, const int32_t*, const float*,
                     const float*, const float*, float*, const float*,
                     const float*, uint32_t, const float*, hipStream_t);
void launch_bi_interp_shepard(const uint32_t*, const int32_t*, const float*,
                              float*, float*, const float*, uint32_t,
                              const float*, const float*, hipStream_t);
void launch_shepard(int, const int32_t*, const float*, const float*,
                    const float*, float*, uint32_t, const uint32_t*,
                    const uint32_t*, const uint32_t*, const uint32_t*,
                    aq_defs, int, hipStream_t, const PartWalk*);
void launch_interactions(int, const int32_t*, const float*, const float*,
                         const float*, const float*, const float*, float*,
                         float*, float*, uint32_t, const uint32_t*,
                         const uint32_t*, const uint32_t*, const uint32_t*,
                         aq_defs, hipStream_t);
void launch_lapp(int, const int32_t*, const float*, const float*,
                 const float*, const float*, float*, uint32_t,
                 const uint32_t*, const uint32_t*, const uint32_t*,
                 const uint32_t*, aq_defs, hipStream_t);
void launch_sensors(int, const int32_t*, const float*, const float*, float*,
                    float*, float*, uint32_t, const uint32_t*,
                    const uint32_t*, const uint32_t*, const uint32_t*,
                    aq_defs, hipStream_t, const float*);
void launch_bi_shepard_compute(int, const int32_t*, const float*,
                               const float*, const float*, const float*,
                               const float*, float*, uint32_t,
                               const uint32_t*, const uint32_t*,
                               const uint32_t*, const uint32_t*, aq_defs,
                               hipStream_t, int, const PartWalk*);
void launch_bi_lapu(int, const int32_t*, const float*, const float*,
                    const float*, const float*, float*, uint32_t,
                    const uint32_t*, const uint32_t*, const uint32_t*,
                    const uint32_t*, aq_defs, hipStream_t,
                    const PartWalk*);
void launch_bi_interp(int, const int32_t*, const float*, const float*,
                      const float*, const float*, float*, uint32_t,
                      const uint32_t*, const uint32_t*, const uint32_t*,
                      const uint32_t*, aq_defs, hipStream_t,
                      const PartWalk*);
void launch_bi_interactions(int, const int32_t*, const float*, const float*,
                            const float*, const float*, const float*,
                            const float*, float*, float*, uint32_t,
                            const uint32_t*, const uint32_t*,
                            const uint32_t*, const uint32_t*, aq_defs,
                            hipStream_t, const uint8_t*, const PartWalk*);
void launch_bi_shepard_compute_cells(int, const int32_t*, const float*,
                                     const float*, const float*,
                                     const float*, const float*, float*,
                                     const CellArgs&, aq_defs, float, int,
                                     hipStream_t);
void launch_bi_lapu_cells(int, const int32_t*, const float*, const float*,
                          const float*, const float*, float*,
                          const CellArgs&, aq_defs, float, hipStream_t);
void launch_bi_interp_cells(int, const int32_t*, const float*, const float*,
                            const float*, const float*, float*,
                            const CellArgs&, aq_defs, float, hipStream_t);
void launch_bi_interactions_cells(int, const int32_t*, const float*,
                                  const float*, const float*, const float*,
                                  const float*, const float*, float*, float*,
                                  const uint8_t*, const CellArgs&, aq_defs,
                                  float, hipStream_t);
void launch_gp_sort(const uint32_t*, uint32_t*, const uint32_t*, uint32_t,
                    hipStream_t);
void launch_gp_mirror(int, const int32_t*, const uint32_t*, const float*,
                      float*, uint32_t*, uint32_t, const float*,
                      const uint32_t*, float, float, hipStream_t);
void launch_gp_interp(int, const int32_t*, const float*, const float*,
                      const float*, const float*, const float*, float*,
                      float*, float*, float*, uint32_t, const uint32_t*,
                      const uint32_t*, const uint32_t*, const uint32_t*,
                      const uint32_t*, aq_defs, hipStream_t);
void launch_gp_renorm(int, const int32_t*, const float*, float*, float*,
                      float*, uint32_t, hipStream_t);
void launch_gp_prelapu(int, const int32_t*, const uint32_t*, const float*,
                       const float*, const float*, float*, float*, uint32_t,
                       int, hipStream_t);
void launch_gp_freeslip(int, const uint32_t*, const int32_t*,
                        const uint32_t*, const float*, float*, uint32_t,
                        const uint32_t*, hipStream_t);
void launch_gp_lapu(int, const int32_t*, const float*, const float*,
                    const float*, const float*, float*, uint32_t,
                    const uint32_t*, const uint32_t*, const uint32_t*,
                    const uint32_t*, aq_defs, hipStream_t);
void launch_gp_preinteractions(int, const uint32_t*, const int32_t*,
                               const uint32_t*, const float*, const float*,
                               const float*, const float*, const float*,
                               const float*, const float*, float*, float*,
                               const float*, uint32_t, const float*,
                               hipStream_t);
void launch_gp_interactions(int, const int32_t*, const float*, const float*,
                            const float*, const float*, const float*, float*,
                            float*, uint32_t, const uint32_t*,
                            const uint32_t*, const uint32_t*,
                            const uint32_t*, aq_defs, hipStream_t);
void launch_bie_interactions(int, const int32_t*, const float*, const float*,
                             const float*, const float*, float*, float*,
                             uint32_t, const uint32_t*, const uint32_t*,
                             const uint32_t*, const uint32_t*, aq_defs,
                             hipStream_t);
void launch_bie_p_boundary(int, const int32_t*, const float*, const float*,
                           const float*, float*, uint32_t, const uint32_t*,
                           const uint32_t*, const uint32_t*, const uint32_t*,
                           aq_defs, hipStream_t);
void launch_bie_rates(int, const int32_t*, const float*, const float*,
                      const float*, const float*, const float*, float*,
                      float*, uint32_t, hipStream_t);
void launch_bie_bounce(int, const int32_t*, const float*, const float*,
                       const float*, const float*, float*, uint32_t,
                       const float*, const uint32_t*, const uint32_t*,
                       const uint32_t*, const uint32_t*, float, float,
                       hipStream_t);
void launch_periodic_wrap(int, float*, uint32_t, const float*, const float*,
                          hipStream_t);
void launch_les_gradu_cells(int, const int32_t*, const float*, const float*,
                            const float*, const float*, float*,
                            const CellArgs&, aq_defs, float, const float*,
                            hipStream_t);
void launch_les_interactions_cells(int, const int32_t*, const float*,
                                   const float*, const float*, const float*,
                                   const float*, float*, const CellArgs&,
                                   aq_defs, float, const float*, hipStream_t);
void launch_elastic_bounce(int, const int32_t*, const float*, const float*,
                           float*, float*, uint32_t, const float*,
                           const float*, const uint32_t*, const uint32_t*,
                           const uint32_t*, const uint32_t*, float, float,
                           float, hipStream_t);
void launch_bie_pst(int, const int32_t*, float*, const float*, const float*,
                    const float*, uint32_t, const uint32_t*, const uint32_t*,
                    const uint32_t*, const uint32_t*, aq_defs, float,
                    hipStream_t);
void launch_mls_entry(int, const int32_t*, const float*, const float*,
                      const float*, float*, const uint32_t*, uint32_t,
                      const uint32_t*, const uint32_t*, const uint32_t*,
                      const uint32_t*, aq_defs, hipStream_t);
void launch_mls_inv(int, const int32_t*, float*, uint32_t, const uint32_t*,
                    hipStream_t);
void launch_lapp_simple(int, const uint32_t*, const int32_t*, float*,
                        const float*, uint32_t, const float*, int,
                        hipStream_t);
void launch_lapp_full(int, const int32_t*, const float*, const float*,
                      const float*, const float*, float*, uint32_t,
                      const uint32_t*, const uint32_t*, const uint32_t*,
                      const uint32_t*, aq_defs, int, hipStream_t);
void launch_lapp_full_mls(int, const int32_t*, const float*, float*,
                          uint32_t, int, hipStream_t);
void launch_lapp_corr(int, const int32_t*, const float*, const float*,
                      const float*, const float*, float*, uint32_t,
                      const uint32_t*, const uint32_t*, const uint32_t*,
                      const uint32_t*, aq_defs, int, hipStream_t);
float support_threshold(float H);
void launch_neighs(int, const int32_t*, uint32_t*, const uint32_t*, uint32_t,
                   const uint32_t*, const uint32_t*, const uint32_t*,
                   const uint32_t*, hipStream_t);
void launch_euler_predictor(int, const float*, const float*, const float*,
                            const float*, const float*, float*, float*,
                            float*, float*, float*, uint32_t, hipStream_t);
void launch_euler_corrector(int, const int32_t*, float*, float*,
                            const float*, float*, const float*, uint32_t,
                            const float*, hipStream_t);

namespace rt {

#define F(k) ((float*)S.ptr(a[k]))
#define I(k) ((int32_t*)S.ptr(a[k]))
#define U(k) ((uint32_t*)S.ptr(a[k]))
#define DB(k) ((double*)S.ptr(a[k]))

static void gather(Sim& S, Var* src, Var* dst, const uint32_t* perm,
                   hipStream_t s)
{
    permute(src->dev, dst->dev, src->type.elem_bytes(), perm, S.N, s);
}

// The gathers (a[k] -> a[k + 1] for even k < 2 nf) of a sort stage in one
// pass over the permutation
static void gather_many(Sim& S, Var* const* a, int nf, const uint32_t* perm,
                        hipStream_t s)
{
    const void* src[8];
    void* dst[8];
    size_t es[8];
    AQ_REQUIRE(nf <= 8, "sort stage: more than 8 fields");
    if (S.no_batch) {
        for (int f = 0; f < nf; f++) gather(S, a[2 * f], a[2 * f + 1], perm, s);
        return;
    }
    for (int f = 0; f < nf; f++) {
        src[f] = a[2 * f]->dev;
        dst[f] = a[2 * f + 1]->dev;
        es[f] = a[2 * f]->type.elem_bytes();
    }
    permute_many(src, dst, es, nf, perm, S.N, s);
}

// __GP_LAPU_MODEL__ names another definition (__GP_LAPU_TAKEDA__ = 1 or
// __GP_LAPU_U0M__ = 2, cfd/GP.xml): follow the chain to the number.
static int gp_lapu_model(Sim& S)
{
    if (!S.defined("__GP_LAPU_MODEL__")) return 2;
    std::string v = S.def("__GP_LAPU_MODEL__");
    for (int k = 0; k < 4 && S.defined(v); k++) v = S.def(v);
    const int m = atoi(v.c_str());
    if (m != 1 && m != 2)
        throw Error(AQ_EUNSUP, "unknown __GP_LAPU_MODEL__ " + v);
    return m;
}

// A preset definition as a float, or a script's own default when no preset
// defines it (the #ifndef defaults of the OpenCL sources).
static float def_or(Sim& S, const char* name, float dflt)
{
    return (S.defined(name) && !S.def(name).empty()) ? S.def_float(name)
                                                     : dflt;
}

static const std::map<std::string, KernelSpec>& table()
{
    static const std::map<std::string, KernelSpec> t = {
        // ---- time integration ------------------------------------------
        { "basic/time_scheme/improved_euler.cl:predictor",
          { { "imove", "r", "u", "dudt", "rho", "drhodt", "r_in", "u_in",
              "dudt_in", "rho_in", "drhodt_in", "dt" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_predictor(S.dims, I(0), F(1), F(2), F(3), F(4), F(5),
                                 F(6), F(7), F(8), F(9), F(10), S.N, F(11),
                                 S.d_status, s);
            } } },
        { "basic/time_scheme/improved_euler.cl:corrector",
          { { "imove", "iset", "r", "u", "dudt", "rho", "drhodt", "dudt_in",
              "drhodt_in", "dt" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_corrector(S.dims, I(0), F(2), F(3), F(4), F(5), F(6),
                                 F(7), F(8), S.N, F(9), S.d_status, s);
            } } },
        { "basic/time_scheme/euler.cl:predictor",
          { { "r", "u", "dudt", "rho", "drhodt", "r_in", "u_in", "dudt_in",
              "rho_in", "drhodt_in" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_euler_predictor(S.dims, F(0), F(1), F(2), F(3), F(4),
                                       F(5), F(6), F(7), F(8), F(9), S.N, s);
            } } },
        { "basic/time_scheme/euler.cl:corrector",
          { { "imove", "iset", "r", "u", "dudt", "rho", "drhodt", "dt" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_euler_corrector(S.dims, I(0), F(2), F(3), F(4), F(5),
                                       F(6), S.N, F(7), s);
            } } },
        { "basic/Domain.cl:entry",
          { { "imove", "r_in", "u_in", "dudt_in", "m", "domain_min",
              "domain_max" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_domain(S.dims, I(0), F(1), F(2), F(3), F(4), S.N, F(5),
                              F(6), s);
            } } },
        // ---- reorder (Sort.cl stage1/stage2 as gathers by id_unsorted) --
        { "basic/Sort.cl:stage1",
          { { "id_in", "id", "iset_in", "iset", "imove_in", "imove", "r_in",
              "r", "normal_in", "normal", "tangent_in", "tangent",
              "id_unsorted" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                gather_many(S, a, 6, U(12), s);
            } } },
        { "basic/Sort.cl:stage2",
          { { "rho_in", "rho", "m_in", "m", "u_in", "u", "dudt", "dudt_in",
              "drhodt", "drhodt_in", "id_unsorted" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                gather_many(S, a, 5, U(10), s);
            } } },
        // ---- basic ----------------------------------------------------------
        { "basic/EOS.cl:entry",
          { { "iset", "imove", "rho", "p", "refd", "cs", "p0" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_eos(U(0), I(1), F(2), F(3), F(4), S.N, F(5), F(6), s);
            } } },
        // Tait EOS (basic/tait.xml; no counterpart script in the reference)
        { "basic/EOS_Tait.cl:entry",
          { { "iset", "imove", "rho", "p", "refd", "cs", "p0", "tait_gamma" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_tait_eos(U(0), I(1), F(2), F(3), F(4), S.N, F(5), F(6),
                                F(7), s);
            } } },
        { "cfd/Boundary/BI/InterpolationShepard_Tait.cl:entry",
          { { "iset", "imove", "shepard", "rho", "p", "refd", "cs", "p0",
              "tait_gamma" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bi_interp_shepard_tait(U(0), I(1), F(2), F(3), F(4),
                                              F(5), S.N, F(6), F(7), F(8), s);
            } } },
        { "basic/Binormal.cl:entry",
          { { "normal", "tangent", "binormal" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_binormal(S.dims, F(0), F(1), F(2), S.N, s);
            } } },
        { "basic/Shepard.cl:entry",
          { { "imove", "r", "rho", "m", "shepard", "icell", "ihoc",
              "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                const PartWalk pw = S.part_walk(a[5], a[6], a[7]);
                launch_shepard(S.dims, I(0), F(1), F(2), F(3), F(4), S.N, U(5),
                               U(6), U(7), S.d_status, S.kdefs, 0, s, &pw);
            } } },
        { "basic/neighs.cl:entry",
          { { "imove", "n_neighs", "neighs_limit", "icell", "ihoc",
              "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_neighs(S.dims, I(0), U(1), U(2), S.N, U(3), U(4), U(5),
                              S.d_status, s);
            } } },
        // ---- cfd --------------------------------------------------------------
        { "cfd/Shepard.cl:entry",
          { { "imove", "r", "rho", "m", "shepard", "icell", "ihoc",
              "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                const PartWalk pw = S.part_walk(a[5], a[6], a[7]);
                launch_shepard(S.dims, I(0), F(1), F(2), F(3), F(4), S.N, U(5),
                               U(6), U(7), S.d_status, S.kdefs, 1, s, &pw);
            } } },
        { "cfd/Interactions.cl:entry",
          { { "imove", "r", "u", "rho", "m", "p", "grad_p", "lap_u", "div_u",
              "icell", "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_interactions(S.dims, I(0), F(1), F(2), F(3), F(4),
                                    F(5), F(6), F(7), F(8), S.N, U(9), U(10),
                                    U(11), S.d_status, S.kdefs, s);
            } } },
        { "cfd/Sensors.cl:entry",
          { { "imove", "r", "m", "u", "rho", "p", "icell", "ihoc",
              "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_sensors(S.dims, I(0), F(1), F(2), F(3), F(4), F(5),
                               S.N, U(6), U(7), U(8), S.d_status, S.kdefs, s,
                               S.periodic_L());
            } } },
        { "cfd/SensorsRenormalization.cl:entry",
          { { "imove", "shepard", "u", "rho", "p" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_sensors_renorm(S.dims, I(0), F(1), F(2), F(3), F(4),
                                      S.N, s);
            } } },
        { "cfd/deltaSPH.cl:lapp",
          { { "imove", "r", "rho", "m", "p", "lap_p", "icell", "ihoc",
              "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_lapp(S.dims, I(0), F(1), F(2), F(3), F(4), F(5), S.N,
                            U(6), U(7), U(8), S.d_status, S.kdefs, s);
            } } },
        { "cfd/deltaSPH.cl:deltaSPH",
          { { "iset", "imove", "rho", "lap_p", "drhodt", "refd", "delta",
              "dt" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_delta_rate(U(0), I(1), F(2), F(3), F(4), F(5), F(6),
                                  S.N, F(7), s);
            } } },
        { "cfd/Rates.cl:entry",
          { { "iset", "imove", "grad_p", "lap_u", "div_u", "dudt", "drhodt",
              "visc_dyn", "g" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_rates(S.dims, U(0), I(1), F(2), F(3), F(4), F(5), F(6),
                             F(7), S.N, F(8), s);
            } } },
        { "cfd/TimeStep.cl:entry",
          { { "imove", "u", "dt_var", "dt", "dt_min", "courant", "dt_Ma",
              "h" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_timestep(S.dims, I(0), F(1), F(2), S.N, F(3), F(4),
                                F(5), F(6), F(7), s);
            } } },
        // ---- plane portals (cfd/portal.xml) ---------------------------------
        { "cfd/Boundary/Portal/Mirror.cl:mirror",
          { { "r", "imirrored", "icell", "portal_in_r", "portal_out_r",
              "portal_n", "r_min", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_portal_mirror(S.dims, F(0), I(1), U(2), S.N, F(3), F(4),
                                     F(5), F(6), U(7), S.kdefs, s);
            } } },
        { "cfd/Boundary/Portal/Mirror.cl:unmirror",
          { { "r", "imirrored", "portal_in_r", "portal_out_r" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_portal_unmirror(S.dims, F(0), I(1), S.N, F(2), F(3), s);
            } } },
        { "cfd/Boundary/Portal/Mirror.cl:teleport",
          { { "r", "portal_in_r", "portal_out_r", "portal_n" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_portal_teleport(S.dims, F(0), S.N, F(1), F(2), F(3), s);
            } } },
        { "cfd/Boundary/Portal/Shepard.cl:entry",
          { { "imove", "imirrored", "r", "rho", "m", "shepard", "icell",
              "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_portal_shepard(S.dims, I(0), I(1), F(2), F(3), F(4),
                                      F(5), S.N, U(6), U(7), U(8), S.d_status,
                                      S.kdefs, support_threshold(S.kdefs.H),
                                      s);
            } } },
        { "cfd/Boundary/Portal/Interactions.cl:entry",
          { { "imove", "imirrored", "r", "u", "rho", "m", "p", "grad_p",
              "lap_u", "div_u", "icell", "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_portal_interactions(
                    S.dims, I(0), I(1), F(2), F(3), F(4), F(5), F(6), F(7),
                    F(8), F(9), S.N, U(10), U(11), U(12), S.d_status, S.kdefs,
                    support_threshold(S.kdefs.H), s);
            } } },
        { "cfd/Boundary/Portal/BoundaryIntegrals.cl:entry",
          { { "imove", "imirrored", "r", "normal", "u", "rho", "m", "p",
              "grad_p", "div_u", "icell", "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_portal_bi(S.dims, I(0), I(1), F(2), F(3), F(4), F(5),
                                 F(6), F(7), F(8), F(9), S.N, U(10), U(11),
                                 U(12), S.d_status, S.kdefs,
                                 support_threshold(S.kdefs.H), s);
            } } },
        { "cfd/Boundary/Portal/deltaSPH.cl:lapp",
          { { "imove", "imirrored", "r", "rho", "m", "p", "lap_p", "icell",
              "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_portal_lapp(S.dims, I(0), I(1), F(2), F(3), F(4), F(5),
                                   F(6), nullptr, S.N, U(7), U(8), U(9),
                                   S.d_status, S.kdefs,
                                   support_threshold(S.kdefs.H), s);
            } } },
        { "cfd/Boundary/Portal/deltaSPH.cl:full",
          { { "imove", "imirrored", "r", "rho", "m", "p", "lap_p_corr",
              "icell", "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_portal_lapp(S.dims, I(0), I(1), F(2), F(3), F(4), F(5),
                                   nullptr, F(6), S.N, U(7), U(8), U(9),
                                   S.d_status, S.kdefs,
                                   support_threshold(S.kdefs.H), s);
            } } },
        { "cfd/Boundary/Portal/deltaSPH.cl:lapp_corr",
          { { "imove", "imirrored", "r", "rho", "m", "lap_p_corr", "lap_p",
              "icell", "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_portal_lapp_corr(S.dims, I(0), I(1), F(2), F(3), F(4),
                                        F(5), F(6), S.N, U(7), U(8), U(9),
                                        S.d_status, S.kdefs,
                                        support_threshold(S.kdefs.H), s);
            } } },
        { "cfd/Boundary/Portal/MLS.cl:entry",
          { { "imove", "imirrored", "r", "rho", "m", "mls", "mls_imove",
              "icell", "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_portal_mls(S.dims, I(0), I(1), F(2), F(3), F(4), F(5),
                                  U(6), S.N, U(7), U(8), U(9), S.d_status,
                                  S.kdefs, support_threshold(S.kdefs.H), s);
            } } },
        // ---- double-precision positions (Presets/precision.xml; no
        // counterpart script in the reference) --------------------------------
        { "basic/time_scheme/improved_euler_fp64.cl:predictor",
          { { "imove", "r", "dr64", "u", "dudt", "rho", "drhodt", "r_in",
              "dr64_in", "u_in", "dudt_in", "rho_in", "drhodt_in", "dt" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_predictor64(S.dims, I(0), F(1), DB(2), F(3), F(4), F(5),
                                   F(6), F(7), DB(8), F(9), F(10), F(11),
                                   F(12), S.N, F(13), S.d_status, s);
            } } },
        { "basic/time_scheme/improved_euler_fp64.cl:corrector",
          { { "imove", "r", "dr64", "u", "dudt", "rho", "drhodt", "dudt_in",
              "drhodt_in", "dt" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_corrector64(S.dims, I(0), F(1), DB(2), F(3), F(4), F(5),
                                   F(6), F(7), F(8), S.N, F(9), S.d_status,
                                   s);
            } } },
        { "basic/Sort_fp64.cl:entry",
          { { "dr64_in", "dr64", "id_unsorted" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                gather(S, a[0], a[1], U(2), s);
            } } },
        { "cfd/FluidMass.cl:entry",
          { { "imove", "m", "m_fluid" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_fluid_mass(I(0), F(1), F(2), S.N, s);
            } } },
        // ---- batched ensembles (Presets/batch.xml; per-set forms of the
        // scalar-parameter kernels, no counterpart script in the reference) --
        { "basic/batch/EOS.cl:entry",
          { { "iset", "imove", "rho", "p", "refd", "cs_set", "p0" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bt_eos(U(0), I(1), F(2), F(3), F(4), F(5), S.N, F(6), s);
            } } },
        { "cfd/batch/Rates.cl:entry",
          { { "iset", "imove", "grad_p", "lap_u", "div_u", "dudt", "drhodt",
              "visc_dyn", "g_set", "exc_amp", "exc_omega", "t_set" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bt_rates(S.dims, U(0), I(1), F(2), F(3), F(4), F(5),
                                F(6), F(7), F(8), F(9), F(10), F(11), S.N, s);
            } } },
        { "cfd/batch/TimeStep.cl:entry",
          { { "iset", "imove", "u", "dt_var", "cs_set", "dt_min", "courant",
              "dt_Ma", "h" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bt_timestep(S.dims, U(0), I(1), F(2), F(3), S.N, F(4),
                                   F(5), F(6), F(7), F(8), s);
            } } },
        { "cfd/batch/TimeStep.cl:min",
          { { "iset", "dt_var", "dt_set" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bt_dt_min(U(0), F(1), S.N, (uint32_t)a[2]->length,
                                 F(2), s);
            } } },
        { "basic/batch/improved_euler.cl:predictor",
          { { "iset", "imove", "r", "u", "dudt", "rho", "drhodt", "r_in",
              "u_in", "dudt_in", "rho_in", "drhodt_in", "dt_set" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bt_predictor(S.dims, U(0), I(1), F(2), F(3), F(4), F(5),
                                    F(6), F(7), F(8), F(9), F(10), F(11),
                                    S.N, F(12), s);
            } } },
        { "basic/batch/improved_euler.cl:corrector",
          { { "iset", "imove", "r", "u", "dudt", "rho", "drhodt", "dudt_in",
              "drhodt_in", "dt_set" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bt_corrector(S.dims, U(0), I(1), F(2), F(3), F(4), F(5),
                                    F(6), F(7), F(8), S.N, F(9), s);
            } } },
        { "cfd/batch/deltaSPH.cl:entry",
          { { "iset", "imove", "rho", "lap_p", "drhodt", "refd", "delta",
              "dt_set" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bt_delta(U(0), I(1), F(2), F(3), F(4), F(5), F(6), S.N,
                                F(7), s);
            } } },
        { "cfd/batch/BI/GradP.cl:entry",
          { { "iset", "imove", "shepard", "lap_u", "dudt", "grad_p",
              "visc_dyn", "refd", "g_set", "exc_amp", "exc_omega", "t_set" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bt_bi_gradp(S.dims, U(0), I(1), F(2), F(3), F(4), F(5),
                                   F(6), F(7), F(8), F(9), F(10), F(11), S.N,
                                   s);
            } } },
        { "cfd/batch/BI/InterpolationShepard.cl:entry",
          { { "iset", "imove", "shepard", "rho", "p", "refd", "cs_set", "p0" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bt_bi_interp_shepard(U(0), I(1), F(2), F(3), F(4), F(5),
                                            F(6), S.N, F(7), s);
            } } },
        { "basic/batch/Time.cl:entry",
          { { "t_set", "dt_set" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bt_time(F(0), F(1), (uint32_t)a[0]->length, s);
            } } },
        // ---- LES closure (Presets/cfd/LES*.xml; no counterpart script in
        // the reference) -------------------------------------------------
        { "cfd/LES/GradU.cl:entry",
          { { "imove", "r", "u", "rho", "m", "grad_u", "icell", "ihoc",
              "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                if (S.clist.usable(a[6], a[7]) && a[8] == S.clist.n_cells)
                    launch_les_gradu_cells(S.dims, I(0), F(1), F(2), F(3),
                                           F(4), F(5), S.cell_args(),
                                           S.kdefs,
                                           support_threshold(S.kdefs.H),
                                           S.periodic_L(), s);
                else if (S.periodic_L())
                    throw Error(AQ_EUNSUP, "periodic box: cfd/LES/GradU.cl "
                                           "needs the cell list");
                else
                    launch_les_gradu(S.dims, I(0), F(1), F(2), F(3), F(4),
                                     F(5), S.N, U(6), U(7), U(8), S.kdefs,
                                     support_threshold(S.kdefs.H), s);
            } } },
        { "cfd/LES/Features.cl:entry",
          { { "imove", "grad_u", "les_in", "hfac" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_les_features(S.dims, I(0), F(1), F(2), S.N,
                                    (float)S.vars.get_host(*a[3])[0], s);
            } } },
        { "cfd/LES/MLP.cl:entry",
          { { "les_in", "les_params", "les_c" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                AQ_REQUIRE(a[0]->length >= 8 * (size_t)S.N &&
                               a[1]->length >= 5313 &&
                               a[2]->length >= S.N,
                           "LES network: les_in (8 N), les_params (5313) or "
                           "les_c (N) too short");
                launch_mlp_forward(F(0), S.N, F(1), F(2), s);
            } } },
        { "cfd/LES/Nut.cl:entry",
          { { "imove", "grad_u", "les_c", "nu_t", "les_cmin" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_les_nut(S.dims, I(0), F(1), F(2), F(3), S.N,
                               S.kdefs.H, 0, 0.f,
                               (float)S.vars.get_host(*a[4])[0], s);
            } } },
        { "cfd/LES/Nut.cl:smagorinsky",
          { { "imove", "grad_u", "nu_t", "les_cs" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_les_nut(S.dims, I(0), F(1), nullptr, F(2), S.N,
                               S.kdefs.H, 1,
                               (float)S.vars.get_host(*a[3])[0], 0.f, s);
            } } },
        { "cfd/LES/Interactions.cl:entry",
          { { "imove", "r", "u", "rho", "m", "nu_t", "dudt", "icell", "ihoc",
              "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                if (S.clist.usable(a[7], a[8]) && a[9] == S.clist.n_cells)
                    launch_les_interactions_cells(
                        S.dims, I(0), F(1), F(2), F(3), F(4), F(5), F(6),
                        S.cell_args(), S.kdefs, support_threshold(S.kdefs.H),
                        S.periodic_L(), s);
                else if (S.periodic_L())
                    throw Error(AQ_EUNSUP, "periodic box: "
                                           "cfd/LES/Interactions.cl needs the "
                                           "cell list");
                else
                    launch_les_interactions(S.dims, I(0), F(1), F(2), F(3),
                                            F(4), F(5), F(6), S.N, U(7), U(8),
                                            U(9), S.kdefs,
                                            support_threshold(S.kdefs.H), s);
            } } },
        // ---- semi-implicit midpoint scheme ----------------------------------
        { "basic/time_scheme/midpoint.cl:predictor",
          { { "r", "u", "dudt", "rho", "drhodt", "r_in", "u_in", "dudt_in",
              "rho_in", "drhodt_in" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_mp_predictor(S.dims, F(0), F(1), F(2), F(3), F(4), F(5),
                                    F(6), F(7), F(8), F(9), S.N, s);
            } } },
        { "basic/time_scheme/midpoint.cl:midpoint",
          { { "imove", "u_in", "u", "dudt", "rho_in", "rho", "drhodt", "dt" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_mp_advance(S.dims, I(0), F(1), F(2), F(3), F(4), F(5),
                                  F(6), S.N, F(7), s);
            } } },
        { "basic/time_scheme/midpoint.cl:midpoint_r",
          { { "imove", "r_in", "r", "u", "dt" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_mp_advance_r(S.dims, I(0), F(1), F(2), F(3), S.N, F(4),
                                    s);
            } } },
        { "basic/time_scheme/midpoint.cl:relax",
          { { "imove", "dudt_in", "dudt", "drhodt_in", "drhodt",
              "relax_midpoint" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_mp_relax(S.dims, I(0), F(1), F(2), F(3), F(4), S.N, F(5),
                                s);
            } } },
        { "basic/time_scheme/midpoint.cl:residuals",
          { { "imove", "m", "u", "dudt_in", "dudt", "rho", "p", "drhodt_in",
              "drhodt", "residual_midpoint" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_mp_residuals(S.dims, I(0), F(1), F(2), F(3), F(4), F(5),
                                    F(6), F(7), F(8), F(9), S.N, s);
            } } },
        { "basic/time_scheme/midpoint.cl:corrector",
          { { "imove", "r_in", "r", "u_in", "u", "dudt", "rho_in", "rho",
              "drhodt", "dt" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_mp_corrector(S.dims, I(0), F(1), F(2), F(3), F(4), F(5),
                                    F(6), F(7), F(8), S.N, F(9), s);
            } } },
        // ---- energy (cfd/energy_kin.xml, cfd/energy.xml) -------------------
        { "cfd/Energy/EnergyKin.cl:entry",
          { { "energy_kin", "imove", "u", "m" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_energy_kin(S.dims, F(0), I(1), F(2), F(3), S.N, s);
            } } },
        { "cfd/Energy/Energy.cl:power",
          { { "energy_dekdt", "energy_depdt", "energy_decdt", "imove", "u",
              "rho", "m", "p", "dudt", "drhodt", "g" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_energy_power(S.dims, F(0), F(1), F(2), I(3), F(4),
                                    F(5), F(6), F(7), F(8), F(9), S.N, F(10),
                                    s);
            } } },
        { "cfd/Energy/Energy.cl:energy",
          { { "energy_ek", "energy_ep", "energy_ec", "iset", "imove", "r", "u",
              "rho", "m", "refd", "g", "cs" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_energy_parts(S.dims, F(0), F(1), F(2), U(3), I(4), F(5),
                                    F(6), F(7), F(8), F(9), S.N, F(10), F(11),
                                    s);
            } } },
        // ---- boundary integrals --------------------------------------------
        { "cfd/Boundary/BI/Shepard.cl:compute",
          { { "imove", "r", "normal", "tangent", "binormal", "m", "shepard",
              "icell", "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                if (S.bi_cells && S.clist.usable(a[7], a[8]) &&
                    a[9] == S.clist.n_cells)
                    launch_bi_shepard_compute_cells(
                        S.dims, I(0), F(1), F(2), F(3), F(4), F(5), F(6),
                        S.cell_args(), S.kdefs, support_threshold(S.kdefs.H),
                        0, s);
                else {
                    const PartWalk pw = S.part_walk(a[7], a[8], a[9]);
                    launch_bi_shepard_compute(S.dims, I(0), F(1), F(2), F(3),
                                              F(4), F(5), F(6), S.N, U(7),
                                              U(8), U(9), S.d_status, S.kdefs,
                                              s, 0, &pw);
                }
            } } },
        { "cfd/Boundary/BI/Shepard.cl:apply",
          { { "imove", "shepard", "grad_p", "lap_u", "div_u" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bi_shepard_apply(S.dims, I(0), F(1), F(2), F(3), F(4),
                                        S.N, s);
            } } },
        { "cfd/Boundary/BI/LapU.cl:freeslip",
          { { "imove", "r", "u", "rho", "m", "lap_u", "icell", "ihoc",
              "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                if (S.bi_cells && S.clist.usable(a[6], a[7]) &&
                    a[8] == S.clist.n_cells)
                    launch_bi_lapu_cells(S.dims, I(0), F(1), F(2), F(3), F(4),
                                         F(5), S.cell_args(), S.kdefs,
                                         support_threshold(S.kdefs.H), s);
                else {
                    const PartWalk pw = S.part_walk(a[6], a[7], a[8]);
                    launch_bi_lapu(S.dims, I(0), F(1), F(2), F(3), F(4), F(5),
                                   S.N, U(6), U(7), U(8), S.d_status, S.kdefs,
                                   s, &pw);
                }
            } } },
        { "cfd/Boundary/BI/GradP.cl:freeslip",
          { { "iset", "imove", "shepard", "lap_u", "dudt", "grad_p",
              "visc_dyn", "refd", "g" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bi_gradp(S.dims, U(0), I(1), F(2), F(3), F(4), F(5),
                                F(6), F(7), S.N, F(8), s);
            } } },
        { "cfd/Boundary/BI/Interpolation.cl:entry",
          { { "imove", "r", "m", "rho", "grad_p", "p", "icell", "ihoc",
              "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                if (S.bi_cells && S.clist.usable(a[6], a[7]) &&
                    a[8] == S.clist.n_cells)
                    launch_bi_interp_cells(S.dims, I(0), F(1), F(2), F(3),
                                           F(4), F(5), S.cell_args(), S.kdefs,
                                           support_threshold(S.kdefs.H), s);
                else {
                    const PartWalk pw = S.part_walk(a[6], a[7], a[8]);
                    launch_bi_interp(S.dims, I(0), F(1), F(2), F(3), F(4),
                                     F(5), S.N, U(6), U(7), U(8), S.d_status,
                                     S.kdefs, s, &pw);
                }
            } } },
        { "cfd/Boundary/BI/InterpolationShepard.cl:entry",
          { { "iset", "imove", "shepard", "rho", "p", "refd", "cs", "p0" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bi_interp_shepard(U(0), I(1), F(2), F(3), F(4), F(5),
                                         S.N, F(6), F(7), s);
            } } },
        // ---- MLS and the delta-SPH Laplacian corrections ------------------
        { "cfd/deltaSPH.cl:simple",
          { { "iset", "imove", "lap_p_corr", "refd", "g" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_lapp_simple(S.dims, U(0), I(1), F(2), F(3), S.N, F(4),
                                   0, s);
            } } },
        { "cfd/deltaSPH.cl:full",
          { { "imove", "r", "rho", "m", "p", "lap_p_corr", "icell", "ihoc",
              "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_lapp_full(S.dims, I(0), F(1), F(2), F(3), F(4), F(5),
                                 S.N, U(6), U(7), U(8), S.d_status, S.kdefs,
                                 0, s);
            } } },
        { "cfd/deltaSPH.cl:full_mls",
          { { "imove", "mls", "lap_p_corr" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_lapp_full_mls(S.dims, I(0), F(1), F(2), S.N, 0, s);
            } } },
        { "cfd/deltaSPH.cl:lapp_corr",
          { { "imove", "r", "rho", "m", "lap_p_corr", "lap_p", "icell",
              "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_lapp_corr(S.dims, I(0), F(1), F(2), F(3), F(4), F(5),
                                 S.N, U(6), U(7), U(8), S.d_status, S.kdefs,
                                 0, s);
            } } },
        { "basic/deltaSPH.cl:simple",
          { { "iset", "imove", "lap_p_corr", "refd", "g" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_lapp_simple(S.dims, U(0), I(1), F(2), F(3), S.N, F(4),
                                   1, s);
            } } },
        { "basic/deltaSPH.cl:full",
          { { "imove", "r", "rho", "m", "p", "lap_p_corr", "icell", "ihoc",
              "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_lapp_full(S.dims, I(0), F(1), F(2), F(3), F(4), F(5),
                                 S.N, U(6), U(7), U(8), S.d_status, S.kdefs,
                                 1, s);
            } } },
        { "basic/deltaSPH.cl:full_mls",
          { { "imove", "mls", "lap_p_corr" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_lapp_full_mls(S.dims, I(0), F(1), F(2), S.N, 1, s);
            } } },
        { "basic/deltaSPH.cl:lapp_corr",
          { { "imove", "r", "rho", "m", "lap_p_corr", "lap_p", "icell",
              "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_lapp_corr(S.dims, I(0), F(1), F(2), F(3), F(4), F(5),
                                 S.N, U(6), U(7), U(8), S.d_status, S.kdefs,
                                 1, s);
            } } },
        { "basic/MLS.cl:entry",
          { { "imove", "r", "rho", "m", "mls", "mls_imove", "icell", "ihoc",
              "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_mls_entry(S.dims, I(0), F(1), F(2), F(3), F(4), U(5),
                                 S.N, U(6), U(7), U(8), S.d_status, S.kdefs, s);
            } } },
        { "basic/MLS.cl:mls_inv",
          { { "imove", "mls", "mls_imove" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_mls_inv(S.dims, I(0), F(1), S.N, U(2), s);
            } } },
        // ---- ghost particles ---------------------------------------------
        { "cfd/Boundary/GP/Sort.cl:entry",
          { { "associations_in", "associations", "id_sorted" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_gp_sort(U(0), U(1), U(2), S.N, s);
            } } },
        { "cfd/Boundary/GP/Mirror.cl:entry",
          { { "imove", "associations", "normal", "r", "gp_icell", "r_min",
              "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_gp_mirror(S.dims, I(0), U(1), F(2), F(3), U(4), S.N,
                                 F(5), U(6), S.kdefs.H, S.def_float("SUPPORT"),
                                 s);
            } } },
        { "cfd/Boundary/GP/Interpolation.cl:entry",
          { { "imove", "r", "m", "rho", "p", "u", "gp_rho", "gp_p", "gp_u",
              "shepard", "gp_icell", "icell", "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_gp_interp(S.dims, I(0), F(1), F(2), F(3), F(4), F(5),
                                 F(6), F(7), F(8), F(9), S.N, U(10), U(11),
                                 U(12), U(13), S.d_status, S.kdefs, s);
            } } },
        { "cfd/Boundary/GP/Renormalization.cl:entry",
          { { "imove", "shepard", "gp_rho", "gp_p", "gp_u" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_gp_renorm(S.dims, I(0), F(1), F(2), F(3), F(4), S.N, s);
            } } },
        { "cfd/Boundary/GP/PreLapU.cl:entry",
          { { "imove", "associations", "normal", "gp_rho", "gp_u", "rho",
              "u" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_gp_prelapu(S.dims, I(0), U(1), F(2), F(3), F(4), F(5),
                                  F(6), S.N, gp_lapu_model(S), s);
            } } },
        { "cfd/Boundary/GP/FreeSlip.cl:entry",
          { { "iset", "imove", "associations", "gp_u", "u", "freeslip_iset" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_gp_freeslip(S.dims, U(0), I(1), U(2), F(3), F(4), S.N,
                                   U(5), s);
            } } },
        { "cfd/Boundary/GP/LapU.cl:entry",
          { { "imove", "r", "u", "rho", "m", "lap_u", "icell", "ihoc",
              "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_gp_lapu(S.dims, I(0), F(1), F(2), F(3), F(4), F(5),
                               S.N, U(6), U(7), U(8), S.d_status, S.kdefs, s);
            } } },
        { "cfd/Boundary/GP/PreInteractions.cl:entry",
          { { "iset", "imove", "associations", "r", "normal", "dudt",
              "lap_u", "gp_p", "gp_u", "rho", "p", "u", "visc_dyn", "g" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_gp_preinteractions(S.dims, U(0), I(1), U(2), F(3),
                                          F(4), F(5), F(6), F(7), F(8), F(9),
                                          F(10), F(11), F(12), S.N, F(13), s);
            } } },
        { "cfd/Boundary/GP/Interactions.cl:entry",
          { { "imove", "r", "u", "rho", "m", "p", "grad_p", "div_u", "icell",
              "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_gp_interactions(S.dims, I(0), F(1), F(2), F(3), F(4),
                                       F(5), F(6), F(7), S.N, U(8), U(9),
                                       U(10), S.d_status, S.kdefs, s);
            } } },
        // ---- conservative boundary integrals ------------------------------
        { "cfd/Boundary/BIe/Interactions.cl:entry",
          { { "imove", "r", "normal", "u", "m", "grad_w_bi", "div_u_bi",
              "icell", "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bie_interactions(S.dims, I(0), F(1), F(2), F(3), F(4),
                                        F(5), F(6), S.N, U(7), U(8), U(9),
                                        S.d_status, S.kdefs, s);
            } } },
        { "cfd/Boundary/BIe/Interactions.cl:p_boundary",
          { { "imove", "r", "m", "rho", "p", "icell", "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bie_p_boundary(S.dims, I(0), F(1), F(2), F(3), F(4),
                                      S.N, U(5), U(6), U(7), S.d_status,
                                      S.kdefs, s);
            } } },
        { "cfd/Boundary/BIe/Rates.cl:entry",
          { { "imove", "rho", "p", "u", "grad_w_bi", "div_u_bi", "grad_p",
              "div_u" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bie_rates(S.dims, I(0), F(1), F(2), F(3), F(4), F(5),
                                 F(6), F(7), S.N, s);
            } } },
        { "cfd/Boundary/BIe/ElasticBounce.cl:entry",
          { { "imove", "r_in", "normal", "m", "u_in", "dudt", "dt", "icell",
              "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bie_bounce(S.dims, I(0), F(1), F(2), F(3), F(4), F(5),
                                  S.N, F(6), U(7), U(8), U(9), S.d_status,
                                  def_or(S, "__DR_FACTOR__", 0.5f),
                                  def_or(S, "__MIN_BOUND_DIST__", 0.f), s);
            } } },
        { "cfd/Boundary/Periodic/Wrap.cl:entry",
          { { "r_in", "periodic_min", "periodic_L" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_periodic_wrap(S.dims, F(0), S.N, F(1), F(2), s);
            } } },
        { "cfd/Boundary/ElasticBounce.cl:entry",
          { { "imove", "r", "normal", "u", "dudt", "dr", "dt", "icell", "ihoc",
              "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_elastic_bounce(S.dims, I(0), F(1), F(2), F(3), F(4),
                                      S.N, F(5), F(6), U(7), U(8), U(9),
                                      S.d_status,
                                      def_or(S, "__DR_FACTOR__", 1.5f),
                                      def_or(S, "__MIN_BOUND_DIST__", 0.3f),
                                      def_or(S, "__ELASTIC_FACTOR__", 0.f), s);
            } } },
        { "cfd/Boundary/BIe/PST.cl:entry",
          { { "imove", "r", "normal", "m", "rho", "icell", "ihoc",
              "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                launch_bie_pst(S.dims, I(0), F(1), F(2), F(3), F(4), S.N, U(5),
                               U(6), U(7), S.d_status, S.kdefs,
                               def_or(S, "__DR_FACTOR__", 0.5f), s);
            } } },
        { "cfd/Boundary/BI/Interactions.cl:entry",
          { { "imove", "r", "normal", "u", "rho", "m", "p", "grad_p", "div_u",
              "icell", "ihoc", "n_cells" },
            [](Sim& S, Var* const* a, hipStream_t s) {
                const uint8_t* near = (S.fusion.active && S.fusion.near)
                                          ? S.fusion.near_buf
                                          : nullptr;
                if (S.bi_cells && S.clist.usable(a[9], a[10]) &&
                    a[11] == S.clist.n_cells)
                    launch_bi_interactions_cells(
                        S.dims, I(0), F(1), F(2), F(3), F(4), F(5), F(6), F(7),
                        F(8), near, S.cell_args(), S.kdefs,
                        support_threshold(S.kdefs.H), s);
                else {
                    const PartWalk pw = S.part_walk(a[9], a[10], a[11]);
                    launch_bi_interactions(S.dims, I(0), F(1), F(2), F(3),
                                           F(4), F(5), F(6), F(7), F(8), S.N,
                                           U(9), U(10), U(11), S.d_status,
                                           S.kdefs, s, near, &pw);
                }
            } } },
    };
    return t;
}

const KernelSpec* find_kernel(const std::string& key)
{
    auto it = table().find(key);
    return it == table().end() ? nullptr : &it->second;
}

std::vector<std::string> registered_kernels()
{
    std::vector<std::string> v;
    for (auto& kv : table()) v.push_back(kv.first);
    return v;
}

std::string kernel_key(const std::string& path, const std::string& entry)
{
    std::string p =
        std::filesystem::path(path).lexically_normal().generic_string();
    const std::string tag = "Scripts/";
    const size_t k = p.rfind(tag);
    if (k != std::string::npos) p = p.substr(k + tag.size());
    return p + ":" + (entry.empty() ? "entry" : entry);
}

}  // namespace rt
}  // namespace aq
