// C exports of the CPU oracle (test infrastructure only; see aqua_oracle.hpp).
// Symbol pattern: orc_<kernel>_<f32|f64>_<2|3>d, arguments in the order of
// the reference kernel signature followed by the preset constants.
#include "aqua_oracle.hpp"

using namespace orc;

#define ORC_EXPORT extern "C" __attribute__((visibility("default")))

// ---- dimension/precision independent ------------------------------------
ORC_EXPORT unsigned orc_radix_key_bits(unsigned long long max_val)
{
    return radix_key_bits(max_val);
}

ORC_EXPORT void orc_stable_sort(usize* keys, usize* perm, usize* inv, usize n,
                                unsigned key_bits)
{
    stable_sort(keys, perm, inv, n, key_bits);
}

ORC_EXPORT void orc_ihoc(const usize* icell_, usize* hoc, usize N,
                         const usize* nc)
{
    ihoc(icell_, hoc, N, nc);
}

ORC_EXPORT void orc_scatter(const void* src, void* dst, size_t esize,
                            const usize* id_sorted, usize N)
{
    scatter(src, dst, esize, id_sorted, N);
}

ORC_EXPORT void orc_gp_sort(const usize* assoc_in, usize* assoc,
                            const usize* id_sorted, usize N)
{
    gp_sort(assoc_in, assoc, id_sorted, N);
}

ORC_EXPORT void orc_neighs_2d(const int* imove, usize* n_neighs, usize limit,
                              usize N, const usize* ic, const usize* hoc,
                              const usize* nc)
{
    neighs<2>(imove, n_neighs, limit, N, LL{ ic, hoc, nc });
}

ORC_EXPORT void orc_neighs_3d(const int* imove, usize* n_neighs, usize limit,
                              usize N, const usize* ic, const usize* hoc,
                              const usize* nc)
{
    neighs<3>(imove, n_neighs, limit, N, LL{ ic, hoc, nc });
}

// ---- per (R, D) ----------------------------------------------------------
#define ORC_DEFINE(R, SR, D)                                                   \
    ORC_EXPORT void orc_minmax_vec_##SR##_##D##d(const R* r, usize N,         \
                                                 R* rmin, R* rmax)            \
    {                                                                          \
        minmax_vec<D, R>(r, N, rmin, rmax);                                    \
    }                                                                          \
    ORC_EXPORT void orc_n_cells_##SR##_##D##d(const R* rmin, const R* rmax,   \
                                              R cell_length, usize* nc)       \
    {                                                                          \
        n_cells<D, R>(rmin, rmax, cell_length, nc);                            \
    }                                                                          \
    ORC_EXPORT void orc_icell_##SR##_##D##d(usize* ic, const R* r, usize N,   \
                                            const R* rmin, R support, R h,    \
                                            const usize* nc)                  \
    {                                                                          \
        icell<D, R>(ic, r, N, rmin, support, h, nc);                           \
    }                                                                          \
    ORC_EXPORT int orc_n_cells_per_##SR##_##D##d(                              \
        const R* rmin, const R* rmax, const R* pmin, const R* L,               \
        R cell_length, usize* nc, R* rmin_out)                                 \
    {                                                                          \
        return n_cells_per<D, R>(rmin, rmax, pmin, L, cell_length, nc,         \
                                 rmin_out)                                     \
                   ? 1                                                         \
                   : 0;                                                        \
    }                                                                          \
    ORC_EXPORT void orc_icell_per_##SR##_##D##d(                               \
        usize* ic, const R* r, usize N, const R* rmin, R support, R h,         \
        const usize* nc, const R* pmin, const R* L)                            \
    {                                                                          \
        icell_per<D, R>(ic, r, N, rmin, support, h, nc, pmin, L);              \
    }                                                                          \
    ORC_EXPORT void orc_periodic_wrap_##SR##_##D##d(R* r, usize N,             \
                                                    const R* pmin, const R* L) \
    {                                                                          \
        periodic_wrap<D, R>(r, N, pmin, L);                                    \
    }                                                                          \
    ORC_EXPORT void orc_predictor_##SR##_##D##d(                               \
        const int* imove, const R* r, const R* u, const R* dudt, const R* rho, \
        const R* drhodt, R* r_in, R* u_in, R* dudt_in, R* rho_in,             \
        R* drhodt_in, usize N, R dt)                                           \
    {                                                                          \
        predictor<D, R>(imove, r, u, dudt, rho, drhodt, r_in, u_in, dudt_in,   \
                        rho_in, drhodt_in, N, dt);                             \
    }                                                                          \
    ORC_EXPORT void orc_corrector_##SR##_##D##d(                               \
        const int* imove, R* r, R* u, const R* dudt, R* rho, const R* drhodt, \
        const R* dudt_in, const R* drhodt_in, usize N, R dt)                   \
    {                                                                          \
        corrector<D, R>(imove, r, u, dudt, rho, drhodt, dudt_in, drhodt_in, N, \
                        dt);                                                   \
    }                                                                          \
    ORC_EXPORT void orc_domain_##SR##_##D##d(int* imove, R* r_in, R* u_in,    \
                                             R* dudt_in, R* m, usize N,       \
                                             const R* dmin, const R* dmax)    \
    {                                                                          \
        domain<D, R>(imove, r_in, u_in, dudt_in, m, N, dmin, dmax);            \
    }                                                                          \
    ORC_EXPORT void orc_eos_##SR##_##D##d(const unsigned* iset,               \
                                          const int* imove, const R* rho,     \
                                          R* p, const R* refd, usize N, R cs, \
                                          R p0)                               \
    {                                                                          \
        eos<R>(iset, imove, rho, p, refd, N, cs, p0);                          \
    }                                                                          \
    ORC_EXPORT void orc_tait_eos_##SR##_##D##d(                                \
        const unsigned* iset, const int* imove, const R* rho, R* p,            \
        const R* refd, usize N, R cs, R p0, R gamma)                           \
    {                                                                          \
        tait_eos<R>(iset, imove, rho, p, refd, N, cs, p0, gamma);              \
    }                                                                          \
    ORC_EXPORT void orc_bi_interp_shepard_tait_##SR##_##D##d(                  \
        const unsigned* iset, const int* imove, const R* shepard_, R* rho,     \
        R* p, const R* refd, usize N, R cs, R p0, R gamma)                     \
    {                                                                          \
        bi_interp_shepard_tait<R>(iset, imove, shepard_, rho, p, refd, N, cs,  \
                                  p0, gamma);                                  \
    }                                                                          \
    ORC_EXPORT void orc_binormal_##SR##_##D##d(const R* normal, R* tangent,   \
                                               R* binormal_, usize N)         \
    {                                                                          \
        binormal<D, R>(normal, tangent, binormal_, N);                         \
    }                                                                          \
    ORC_EXPORT void orc_portal_mirror_##SR##_##D##d(                           \
        R* r, int* imirrored, usize* icell, usize N, const R* in_r,            \
        const R* out_r, const R* n, const R* r_min, const usize* nc, R H)      \
    {                                                                          \
        portal_mirror<D, R>(r, imirrored, icell, N, in_r, out_r, n, r_min, nc, \
                            H);                                                \
    }                                                                          \
    ORC_EXPORT void orc_portal_unmirror_##SR##_##D##d(                         \
        R* r, const int* imirrored, usize N, const R* in_r, const R* out_r)    \
    {                                                                          \
        portal_unmirror<D, R>(r, imirrored, N, in_r, out_r);                   \
    }                                                                          \
    ORC_EXPORT void orc_portal_teleport_##SR##_##D##d(                         \
        R* r, usize N, const R* in_r, const R* out_r, const R* n)              \
    {                                                                          \
        portal_teleport<D, R>(r, N, in_r, out_r, n);                           \
    }                                                                          \
    ORC_EXPORT void orc_portal_shepard_##SR##_##D##d(                          \
        const int* imove, const int* imirrored, const R* r, const R* rho,      \
        const R* m, R* shep, usize N, const usize* ic, const usize* hoc,       \
        const usize* nc, R H, R CONW)                                          \
    {                                                                          \
        portal_shepard<D, R>(imove, imirrored, r, rho, m, shep, N,             \
                             LL{ ic, hoc, nc }, H, CONW);                      \
    }                                                                          \
    ORC_EXPORT void orc_portal_interactions_##SR##_##D##d(                     \
        const int* imove, const int* imirrored, const R* r, const R* u,        \
        const R* rho, const R* m, const R* p, R* grad_p, R* lap_u, R* div_u,   \
        usize N, const usize* ic, const usize* hoc, const usize* nc, R H,      \
        R CONF, int lap)                                                       \
    {                                                                          \
        portal_interactions<D, R>(imove, imirrored, r, u, rho, m, p, grad_p,   \
                                  lap_u, div_u, N, LL{ ic, hoc, nc }, H, CONF, \
                                  lap);                                        \
    }                                                                          \
    ORC_EXPORT void orc_portal_bi_##SR##_##D##d(                               \
        const int* imove, const int* imirrored, const R* r, const R* normal,   \
        const R* u, const R* rho, const R* m, const R* p, R* grad_p,           \
        R* div_u, usize N, const usize* ic, const usize* hoc, const usize* nc, \
        R H, R CONW)                                                           \
    {                                                                          \
        portal_bi<D, R>(imove, imirrored, r, normal, u, rho, m, p, grad_p,     \
                        div_u, N, LL{ ic, hoc, nc }, H, CONW);                 \
    }                                                                          \
    ORC_EXPORT void orc_portal_lapp_##SR##_##D##d(                             \
        const int* imove, const int* imirrored, const R* r, const R* rho,      \
        const R* m, const R* p, R* lap_p, R* lap_p_corr, usize N,              \
        const usize* ic, const usize* hoc, const usize* nc, R H, R CONF)       \
    {                                                                          \
        portal_lapp<D, R>(imove, imirrored, r, rho, m, p, lap_p, lap_p_corr,   \
                          N, LL{ ic, hoc, nc }, H, CONF);                      \
    }                                                                          \
    ORC_EXPORT void orc_portal_lapp_corr_##SR##_##D##d(                        \
        const int* imove, const int* imirrored, const R* r, const R* rho,      \
        const R* m, const R* lap_p_corr, R* lap_p, usize N, const usize* ic,   \
        const usize* hoc, const usize* nc, R H, R CONF)                        \
    {                                                                          \
        portal_lapp_corr<D, R>(imove, imirrored, r, rho, m, lap_p_corr,        \
                               lap_p, N, LL{ ic, hoc, nc }, H, CONF);          \
    }                                                                          \
    ORC_EXPORT void orc_portal_mls_##SR##_##D##d(                              \
        const int* imove, const int* imirrored, const R* r, const R* rho,      \
        const R* m, R* mls, usize N, unsigned mls_imove, const usize* ic,      \
        const usize* hoc, const usize* nc, R H, R CONF)                        \
    {                                                                          \
        portal_mls<D, R>(imove, imirrored, r, rho, m, mls, N, mls_imove,       \
                         LL{ ic, hoc, nc }, H, CONF);                          \
    }                                                                          \
    ORC_EXPORT void orc_midpoint_predictor_##SR##_##D##d(                      \
        const R* r, const R* u, const R* dudt, const R* rho, const R* drhodt,  \
        R* r_in, R* u_in, R* dudt_in, R* rho_in, R* drhodt_in, usize N)        \
    {                                                                          \
        midpoint_predictor<D, R>(r, u, dudt, rho, drhodt, r_in, u_in, dudt_in, \
                                 rho_in, drhodt_in, N);                        \
    }                                                                          \
    ORC_EXPORT void orc_midpoint_advance_##SR##_##D##d(                        \
        const int* imove, const R* u_in, R* u, const R* dudt, const R* rho_in, \
        R* rho, const R* drhodt, usize N, R dt)                                \
    {                                                                          \
        midpoint_advance<D, R>(imove, u_in, u, dudt, rho_in, rho, drhodt, N,   \
                               dt);                                            \
    }                                                                          \
    ORC_EXPORT void orc_midpoint_r_##SR##_##D##d(                              \
        const int* imove, const R* r_in, R* r, const R* u, usize N, R dt)      \
    {                                                                          \
        midpoint_r<D, R>(imove, r_in, r, u, N, dt);                            \
    }                                                                          \
    ORC_EXPORT void orc_midpoint_relax_##SR##_##D##d(                          \
        const int* imove, const R* dudt_in, R* dudt, const R* drhodt_in,       \
        R* drhodt, usize N, R f)                                               \
    {                                                                          \
        midpoint_relax<D, R>(imove, dudt_in, dudt, drhodt_in, drhodt, N, f);   \
    }                                                                          \
    ORC_EXPORT void orc_midpoint_residuals_##SR##_##D##d(                      \
        const int* imove, const R* m, const R* u, const R* dudt_in,            \
        const R* dudt, const R* rho, const R* p, const R* drhodt_in,           \
        const R* drhodt, R* res, usize N)                                      \
    {                                                                          \
        midpoint_residuals<D, R>(imove, m, u, dudt_in, dudt, rho, p,           \
                                 drhodt_in, drhodt, res, N);                   \
    }                                                                          \
    ORC_EXPORT void orc_midpoint_corrector_##SR##_##D##d(                      \
        const int* imove, const R* r_in, R* r, const R* u_in, R* u,            \
        const R* dudt, const R* rho_in, R* rho, const R* drhodt, usize N,      \
        R dt)                                                                  \
    {                                                                          \
        midpoint_corrector<D, R>(imove, r_in, r, u_in, u, dudt, rho_in, rho,   \
                                 drhodt, N, dt);                               \
    }                                                                          \
    ORC_EXPORT void orc_energy_kin_##SR##_##D##d(                              \
        R* ek, const int* imove, const R* u, const R* m, usize N)              \
    {                                                                          \
        energy_kin<D, R>(ek, imove, u, m, N);                                  \
    }                                                                          \
    ORC_EXPORT void orc_energy_power_##SR##_##D##d(                            \
        R* dek, R* dep, R* dec, const int* imove, const R* u, const R* rho,    \
        const R* m, const R* p, const R* dudt, const R* drhodt, usize N,       \
        const R* g)                                                            \
    {                                                                          \
        energy_power<D, R>(dek, dep, dec, imove, u, rho, m, p, dudt, drhodt,   \
                           N, g);                                              \
    }                                                                          \
    ORC_EXPORT void orc_energy_parts_##SR##_##D##d(                            \
        R* ek, R* ep, R* ec, const unsigned* iset, const int* imove,           \
        const R* r, const R* u, const R* rho, const R* m, const R* refd,       \
        usize N, const R* g, R cs)                                             \
    {                                                                          \
        energy_parts<D, R>(ek, ep, ec, iset, imove, r, u, rho, m, refd, N, g,  \
                           cs);                                                \
    }                                                                          \
    ORC_EXPORT void orc_rates_##SR##_##D##d(                                   \
        const unsigned* iset, const int* imove, const R* rho,                  \
        const R* grad_p, const R* lap_u, const R* div_u, R* dudt, R* drhodt,   \
        const R* visc_dyn, usize N, const R* g)                                \
    {                                                                          \
        rates<D, R>(iset, imove, rho, grad_p, lap_u, div_u, dudt, drhodt,      \
                    visc_dyn, N, g);                                           \
    }                                                                          \
    ORC_EXPORT void orc_delta_rate_##SR##_##D##d(                              \
        const unsigned* iset, const int* imove, const R* rho, const R* lap_p,  \
        R* drhodt, const R* refd, const R* delta, usize N, R dt)               \
    {                                                                          \
        delta_rate<R>(iset, imove, rho, lap_p, drhodt, refd, delta, N, dt);    \
    }                                                                          \
    ORC_EXPORT void orc_timestep_##SR##_##D##d(const int* imove, const R* u,  \
                                               R* dt_var, usize N, R dt,      \
                                               R dt_min, R courant, R dt_Ma,  \
                                               R h)                           \
    {                                                                          \
        timestep<D, R>(imove, u, dt_var, N, dt, dt_min, courant, dt_Ma, h);    \
    }                                                                          \
    ORC_EXPORT void orc_bi_shepard_apply_##SR##_##D##d(                        \
        const int* imove, const R* shepard_, R* grad_p, R* lap_u, R* div_u,   \
        usize N)                                                               \
    {                                                                          \
        bi_shepard_apply<D, R>(imove, shepard_, grad_p, lap_u, div_u, N);      \
    }                                                                          \
    ORC_EXPORT void orc_bi_gradp_##SR##_##D##d(                                \
        const unsigned* iset, const int* imove, const R* shepard_,             \
        const R* rho, const R* lap_u, const R* dudt, R* grad_p,                \
        const R* visc_dyn, const R* refd, usize N, const R* g)                 \
    {                                                                          \
        bi_gradp<D, R>(iset, imove, shepard_, rho, lap_u, dudt, grad_p,        \
                       visc_dyn, refd, N, g);                                  \
    }                                                                          \
    ORC_EXPORT void orc_bi_interp_shepard_##SR##_##D##d(                       \
        const unsigned* iset, const int* imove, const R* shepard_, R* rho,     \
        R* p, const R* refd, usize N, R cs, R p0)                              \
    {                                                                          \
        bi_interp_shepard<R>(iset, imove, shepard_, rho, p, refd, N, cs, p0);  \
    }                                                                          \
    ORC_EXPORT void orc_sensors_renorm_##SR##_##D##d(                          \
        const int* imove, const R* shepard_, R* u, R* rho, R* p, usize N)      \
    {                                                                          \
        sensors_renorm<D, R>(imove, shepard_, u, rho, p, N);                   \
    }                                                                          \
    ORC_EXPORT void orc_shepard_##SR##_##D##d(                                 \
        const int* imove, const R* r, const R* rho, const R* m, R* shep,       \
        usize N, const usize* ic, const usize* hoc, const usize* nc, R H,      \
        R CONW, int cfd_excl)                                                  \
    {                                                                          \
        shepard<D, R>(imove, r, rho, m, shep, N, LL{ ic, hoc, nc }, H, CONW,   \
                      cfd_excl);                                               \
    }                                                                          \
    ORC_EXPORT void orc_interactions_##SR##_##D##d(                            \
        const int* imove, const R* r, const R* u, const R* rho, const R* m,    \
        const R* p, R* grad_p, R* lap_u, R* div_u, usize N, const usize* ic,   \
        const usize* hoc, const usize* nc, R H, R CONF, int lap_form)          \
    {                                                                          \
        interactions<D, R>(imove, r, u, rho, m, p, grad_p, lap_u, div_u, N,    \
                           LL{ ic, hoc, nc }, H, CONF, lap_form);              \
    }                                                                          \
    ORC_EXPORT void orc_lapp_##SR##_##D##d(                                    \
        const int* imove, const R* r, const R* rho, const R* m, const R* p,    \
        R* lap_p, usize N, const usize* ic, const usize* hoc, const usize* nc, \
        R H, R CONF)                                                           \
    {                                                                          \
        lapp<D, R>(imove, r, rho, m, p, lap_p, N, LL{ ic, hoc, nc }, H, CONF); \
    }                                                                          \
    ORC_EXPORT void orc_les_gradu_##SR##_##D##d(                               \
        const int* imove, const R* r, const R* u, const R* rho, const R* m,    \
        R* grad_u, usize N, const usize* ic, const usize* hoc,                 \
        const usize* nc, R H, R CONF)                                          \
    {                                                                          \
        les_gradu<D, R>(imove, r, u, rho, m, grad_u, N, LL{ ic, hoc, nc }, H,  \
                        CONF);                                                 \
    }                                                                          \
    ORC_EXPORT void orc_les_interactions_##SR##_##D##d(                        \
        const int* imove, const R* r, const R* u, const R* rho, const R* m,    \
        const R* nu_t, R* dudt, usize N, const usize* ic, const usize* hoc,    \
        const usize* nc, R H, R CONF)                                          \
    {                                                                          \
        les_interactions<D, R>(imove, r, u, rho, m, nu_t, dudt, N,             \
                               LL{ ic, hoc, nc }, H, CONF);                    \
    }                                                                          \
    ORC_EXPORT void orc_interactions_kahan_##SR##_##D##d(                      \
        const int* imove, const R* r, const R* u, const R* rho, const R* m,    \
        const R* p, R* grad_p, R* lap_u, R* div_u, usize N, const usize* ic,   \
        const usize* hoc, const usize* nc, R H, R CONF, int lap_form)          \
    {                                                                          \
        interactions<D, R>(imove, r, u, rho, m, p, grad_p, lap_u, div_u, N,    \
                           LL{ ic, hoc, nc }, H, CONF, lap_form, true);        \
    }                                                                          \
    ORC_EXPORT void orc_lapp_kahan_##SR##_##D##d(                              \
        const int* imove, const R* r, const R* rho, const R* m, const R* p,    \
        R* lap_p, usize N, const usize* ic, const usize* hoc, const usize* nc, \
        R H, R CONF)                                                           \
    {                                                                          \
        lapp<D, R>(imove, r, rho, m, p, lap_p, N, LL{ ic, hoc, nc }, H, CONF,  \
                   true);                                                      \
    }                                                                          \
    ORC_EXPORT void orc_sensors_##SR##_##D##d(                                 \
        const int* imove, const R* r, const R* m, R* u, R* rho, R* p, usize N, \
        const usize* ic, const usize* hoc, const usize* nc, R H, R CONW)       \
    {                                                                          \
        sensors<D, R>(imove, r, m, u, rho, p, N, LL{ ic, hoc, nc }, H, CONW);  \
    }                                                                          \
    ORC_EXPORT void orc_bi_shepard_compute_##SR##_##D##d(                      \
        const int* imove, const R* r, const R* normal, const R* tangent,       \
        const R* binormal_, const R* m, R* shep, usize N, const usize* ic,     \
        const usize* hoc, const usize* nc, R H, R CONW, R DIMS)                \
    {                                                                          \
        bi_shepard_compute<D, R>(imove, r, normal, tangent, binormal_, m,      \
                                 shep, N, LL{ ic, hoc, nc }, H, CONW, DIMS);   \
    }                                                                          \
    ORC_EXPORT void orc_bi_lapu_##SR##_##D##d(                                 \
        const int* imove, const R* r, const R* u, const R* rho, const R* m,    \
        R* lap_u, usize N, const usize* ic, const usize* hoc, const usize* nc, \
        R H, R CONF, int lap_form)                                             \
    {                                                                          \
        bi_lapu<D, R>(imove, r, u, rho, m, lap_u, N, LL{ ic, hoc, nc }, H,     \
                      CONF, lap_form);                                         \
    }                                                                          \
    ORC_EXPORT void orc_bi_interp_##SR##_##D##d(                               \
        const int* imove, const R* r, const R* m, const R* rho,                \
        const R* grad_p, R* p, usize N, const usize* ic, const usize* hoc,     \
        const usize* nc, R H, R CONW)                                          \
    {                                                                          \
        bi_interp<D, R>(imove, r, m, rho, grad_p, p, N, LL{ ic, hoc, nc }, H,  \
                        CONW);                                                 \
    }                                                                          \
    ORC_EXPORT void orc_bi_interactions_##SR##_##D##d(                         \
        const int* imove, const R* r, const R* normal, const R* u,             \
        const R* rho, const R* m, const R* p, R* grad_p, R* div_u, usize N,    \
        const usize* ic, const usize* hoc, const usize* nc, R H, R CONW)       \
    {                                                                          \
        bi_interactions<D, R>(imove, r, normal, u, rho, m, p, grad_p, div_u,   \
                              N, LL{ ic, hoc, nc }, H, CONW);                  \
    }                                                                          \
    ORC_EXPORT void orc_gp_mirror_##SR##_##D##d(                               \
        const int* imove, const usize* assoc, const R* normal, R* r,           \
        usize* gp_icell, usize N, const R* rmin, const usize* nc, R H,         \
        R SUPPORT)                                                             \
    {                                                                          \
        gp_mirror<D, R>(imove, assoc, normal, r, gp_icell, N, rmin, nc, H,     \
                        SUPPORT);                                              \
    }                                                                          \
    ORC_EXPORT void orc_gp_interp_##SR##_##D##d(                               \
        const int* imove, const R* r, const R* m, const R* rho, const R* p,    \
        const R* u, R* gp_rho, R* gp_p, R* gp_u, R* shepard, usize N,          \
        const usize* gp_icell, const usize* ic, const usize* hoc,              \
        const usize* nc, R H, R CONW)                                          \
    {                                                                          \
        gp_interp<D, R>(imove, r, m, rho, p, u, gp_rho, gp_p, gp_u, shepard,   \
                        N, gp_icell, LL{ ic, hoc, nc }, H, CONW);              \
    }                                                                          \
    ORC_EXPORT void orc_gp_renorm_##SR##_##D##d(const int* imove,             \
                                                const R* shepard, R* gp_rho,  \
                                                R* gp_p, R* gp_u, usize N)    \
    {                                                                          \
        gp_renorm<D, R>(imove, shepard, gp_rho, gp_p, gp_u, N);                \
    }                                                                          \
    ORC_EXPORT void orc_gp_prelapu_##SR##_##D##d(                              \
        const int* imove, const usize* assoc, const R* normal,                 \
        const R* gp_rho, const R* gp_u, R* rho, R* u, usize N, int model)      \
    {                                                                          \
        gp_prelapu<D, R>(imove, assoc, normal, gp_rho, gp_u, rho, u, N,        \
                         model);                                               \
    }                                                                          \
    ORC_EXPORT void orc_gp_freeslip_##SR##_##D##d(                             \
        const unsigned* iset, const int* imove, const usize* assoc,            \
        const R* gp_u, R* u, usize N, unsigned fs)                             \
    {                                                                          \
        gp_freeslip<D, R>(iset, imove, assoc, gp_u, u, N, fs);                 \
    }                                                                          \
    ORC_EXPORT void orc_gp_lapu_##SR##_##D##d(                                 \
        const int* imove, const R* r, const R* u, const R* rho, const R* m,    \
        R* lap_u, usize N, const usize* ic, const usize* hoc,                  \
        const usize* nc, R H, R CONF, int lap_form)                            \
    {                                                                          \
        gp_lapu<D, R>(imove, r, u, rho, m, lap_u, N, LL{ ic, hoc, nc }, H,     \
                      CONF, lap_form);                                         \
    }                                                                          \
    ORC_EXPORT void orc_gp_preinteractions_##SR##_##D##d(                      \
        const unsigned* iset, const int* imove, const usize* assoc,            \
        const R* r, const R* normal, const R* dudt, const R* lap_u,            \
        const R* gp_p, const R* gp_u, const R* rho, R* p, R* u,                \
        const R* visc_dyn, usize N, const R* g)                                \
    {                                                                          \
        gp_preinteractions<D, R>(iset, imove, assoc, r, normal, dudt, lap_u,   \
                                 gp_p, gp_u, rho, p, u, visc_dyn, N, g);       \
    }                                                                          \
    ORC_EXPORT void orc_gp_interactions_##SR##_##D##d(                         \
        const int* imove, const R* r, const R* u, const R* rho, const R* m,    \
        const R* p, R* grad_p, R* div_u, usize N, const usize* ic,             \
        const usize* hoc, const usize* nc, R H, R CONF)                        \
    {                                                                          \
        gp_interactions<D, R>(imove, r, u, rho, m, p, grad_p, div_u, N,        \
                              LL{ ic, hoc, nc }, H, CONF);                     \
    }                                                                          \
    ORC_EXPORT void orc_bie_interactions_##SR##_##D##d(                        \
        const int* imove, const R* r, const R* normal, const R* u,             \
        const R* m, R* grad_w_bi, R* div_u_bi, usize N, const usize* ic,       \
        const usize* hoc, const usize* nc, R H, R CONW)                        \
    {                                                                          \
        bie_interactions<D, R>(imove, r, normal, u, m, grad_w_bi, div_u_bi, N, \
                               LL{ ic, hoc, nc }, H, CONW);                    \
    }                                                                          \
    ORC_EXPORT void orc_bie_p_boundary_##SR##_##D##d(                          \
        const int* imove, const R* r, const R* m, const R* rho, R* p, usize N, \
        const usize* ic, const usize* hoc, const usize* nc, R H, R CONW)       \
    {                                                                          \
        bie_p_boundary<D, R>(imove, r, m, rho, p, N, LL{ ic, hoc, nc }, H,     \
                             CONW);                                            \
    }                                                                          \
    ORC_EXPORT void orc_bie_rates_##SR##_##D##d(                               \
        const int* imove, const R* rho, const R* p, const R* u,                \
        const R* grad_w_bi, const R* div_u_bi, R* grad_p, R* div_u, usize N)   \
    {                                                                          \
        bie_rates<D, R>(imove, rho, p, u, grad_w_bi, div_u_bi, grad_p, div_u,  \
                        N);                                                    \
    }                                                                          \
    ORC_EXPORT void orc_bie_elastic_bounce_##SR##_##D##d(                      \
        const int* imove, const R* r_in, const R* normal, const R* m,          \
        const R* u_in, R* dudt, usize N, R dt, const usize* ic,                \
        const usize* hoc, const usize* nc, R DR_FACTOR, R MIN_BOUND_DIST)      \
    {                                                                          \
        bie_elastic_bounce<D, R>(imove, r_in, normal, m, u_in, dudt, N, dt,    \
                                 LL{ ic, hoc, nc }, DR_FACTOR,                 \
                                 MIN_BOUND_DIST);                              \
    }                                                                          \
    ORC_EXPORT void orc_elastic_bounce_##SR##_##D##d(                          \
        const int* imove, const R* r, const R* normal, R* u, R* dudt, usize N, \
        R dr, R dt, const usize* ic, const usize* hoc, const usize* nc,        \
        R DR_FACTOR, R MIN_BOUND_DIST, R ELASTIC_FACTOR)                       \
    {                                                                          \
        elastic_bounce<D, R>(imove, r, normal, u, dudt, N, dr, dt,             \
                             LL{ ic, hoc, nc }, DR_FACTOR, MIN_BOUND_DIST,     \
                             ELASTIC_FACTOR);                                  \
    }                                                                          \
    ORC_EXPORT void orc_bie_pst_##SR##_##D##d(                                 \
        const int* imove, R* r, const R* normal, const R* m, const R* rho,     \
        usize N, const usize* ic, const usize* hoc, const usize* nc, R DIMS,   \
        R DR_FACTOR)                                                           \
    {                                                                          \
        bie_pst<D, R>(imove, r, normal, m, rho, N, LL{ ic, hoc, nc }, DIMS,    \
                      DR_FACTOR);                                              \
    }                                                                          \
    ORC_EXPORT void orc_mls_entry_##SR##_##D##d(                               \
        const int* imove, const R* r, const R* rho, const R* m, R* mls,        \
        usize N, unsigned mls_imove, const usize* ic, const usize* hoc,        \
        const usize* nc, R H, R CONF)                                          \
    {                                                                          \
        mls_entry<D, R>(imove, r, rho, m, mls, N, mls_imove,                   \
                        LL{ ic, hoc, nc }, H, CONF);                           \
    }                                                                          \
    ORC_EXPORT void orc_mls_inv_##SR##_##D##d(const int* imove, R* mls,       \
                                              usize N, unsigned mls_imove)    \
    {                                                                          \
        mls_inv<D, R>(imove, mls, N, mls_imove);                               \
    }                                                                          \
    ORC_EXPORT void orc_lapp_simple_##SR##_##D##d(                             \
        const unsigned* iset, const int* imove, R* lap_p_corr, const R* refd,  \
        usize N, const R* g)                                                   \
    {                                                                          \
        lapp_simple<D, R>(iset, imove, lap_p_corr, refd, N, g);                \
    }                                                                          \
    ORC_EXPORT void orc_lapp_full_##SR##_##D##d(                               \
        const int* imove, const R* r, const R* rho, const R* m, const R* p,    \
        R* lap_p_corr, usize N, const usize* ic, const usize* hoc,             \
        const usize* nc, R H, R CONF)                                          \
    {                                                                          \
        lapp_full<D, R>(imove, r, rho, m, p, lap_p_corr, N, LL{ ic, hoc, nc }, \
                        H, CONF);                                              \
    }                                                                          \
    ORC_EXPORT void orc_lapp_full_mls_##SR##_##D##d(                           \
        const int* imove, const R* mls, R* lap_p_corr, usize N)                \
    {                                                                          \
        lapp_full_mls<D, R>(imove, mls, lap_p_corr, N);                        \
    }                                                                          \
    ORC_EXPORT void orc_lapp_corr_##SR##_##D##d(                               \
        const int* imove, const R* r, const R* rho, const R* m,                \
        const R* lap_p_corr, R* lap_p, usize N, const usize* ic,               \
        const usize* hoc, const usize* nc, R H, R CONF)                        \
    {                                                                          \
        lapp_corr<D, R>(imove, r, rho, m, lap_p_corr, lap_p, N,                \
                        LL{ ic, hoc, nc }, H, CONF);                           \
    }

ORC_DEFINE(float, f32, 2)
ORC_DEFINE(float, f32, 3)

// precision mode: float state with double positions (r + dr64)
#define ORC_DEFINE64(D)                                                        \
    ORC_EXPORT void orc_predictor64_##D##d(                                    \
        const int* imove, const float* r, const double* dr64, const float* u,  \
        const float* dudt, const float* rho, const float* drhodt,             \
        float* r_in, double* dr64_in, float* u_in, float* dudt_in,            \
        float* rho_in, float* drhodt_in, usize N, float dt)                    \
    {                                                                          \
        predictor64<D>(imove, r, dr64, u, dudt, rho, drhodt, r_in, dr64_in,    \
                       u_in, dudt_in, rho_in, drhodt_in, N, dt);               \
    }                                                                          \
    ORC_EXPORT void orc_corrector64_##D##d(                                    \
        const int* imove, float* r, double* dr64, float* u, const float* dudt, \
        float* rho, const float* drhodt, const float* dudt_in,                 \
        const float* drhodt_in, usize N, float dt)                             \
    {                                                                          \
        corrector64<D>(imove, r, dr64, u, dudt, rho, drhodt, dudt_in,          \
                       drhodt_in, N, dt);                                      \
    }
ORC_DEFINE64(2)
ORC_DEFINE64(3)

// the compensated sum of x[0..n) in order: out[0] = s, out[1] = c
ORC_EXPORT void orc_nsum_f32(const float* x, usize n, float* out)
{
    float s = 0.f, c = 0.f;
    for (usize i = 0; i < n; i++) nsum(s, c, x[i]);
    out[0] = s;
    out[1] = c;
}
ORC_DEFINE(double, f64, 2)
ORC_DEFINE(double, f64, 3)

// ---- kernel-function probes (FP64 instantiation of the FP32 constants) ---
ORC_EXPORT double orc_kernel_w(int dims, double q)
{
    return dims == 2 ? kernelW<2, double>(q) : kernelW<3, double>(q);
}
ORC_EXPORT double orc_kernel_f(int dims, double q)
{
    return dims == 2 ? kernelF<2, double>(q) : kernelF<3, double>(q);
}
ORC_EXPORT double orc_kernel_sp(int dims, double q)
{
    return dims == 2 ? kernelS_P<2, double>(q) : kernelS_P<3, double>(q);
}
ORC_EXPORT double orc_kernel_sd(int dims, double d, double t, double b,
                                double s)
{
    return dims == 2 ? kernelS_D<2, double>(d, t, b, s)
                     : kernelS_D<3, double>(d, t, b, s);
}

// ---- the float atan / acospi sequence of the boundary-integral terms -----
// op 0: atan, op 1: acospi.  y[k] = the float sequence at x[k].
ORC_EXPORT void orc_clmath_f32(int op, const float* x, float* y, usize n)
{
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < (long long)n; k++)
        y[k] = op == 0 ? atan_cl(x[k]) : acospi_cl(x[k]);
}

// Largest error, in units of the float spacing at the exact value (libm in
// double), of the float sequence over the float bit patterns lo, lo+stride,
// ... <= hi (op 0: atan, op 1: acospi).  out[0] = max ulp, out[1] = the
// argument where it occurs.
ORC_EXPORT void orc_clmath_ulp(int op, unsigned lo, unsigned hi,
                               unsigned stride, double* out)
{
    const long long n = ((long long)hi - (long long)lo) / stride + 1;
    double best = 0.0, arg = 0.0;
#pragma omp parallel
    {
        double b = 0.0, a = 0.0;
#pragma omp for schedule(static)
        for (long long k = 0; k < n; k++) {
            const unsigned bits = lo + (unsigned)(k * stride);
            float x;
            std::memcpy(&x, &bits, 4);
            if (!std::isfinite(x)) continue;
            const double ex = op == 0 ? std::atan((double)x)
                                      : std::acos((double)x) / M_PI;
            const double got = op == 0 ? atan_cl(x) : acospi_cl(x);
            const double ax = std::fabs(ex);
            int e = 0;
            std::frexp(ax, &e);  // ax in [2^(e-1), 2^e)
            const double ulp = std::ldexp(1.0, std::max(e - 24, -149));
            const double err = std::fabs(got - ex) / ulp;
            if (err > b) {
                b = err;
                a = x;
            }
        }
#pragma omp critical
        if (b > best) {
            best = b;
            arg = a;
        }
    }
    out[0] = best;
    out[1] = arg;
}

// The periodic box of the oracle's walks: lengths along x, y, z (0: that
// axis is not periodic; all 0: no periodic box).
ORC_EXPORT void orc_set_periodic(double Lx, double Ly, double Lz)
{
    orc::g_perL[0] = Lx;
    orc::g_perL[1] = Ly;
    orc::g_perL[2] = Lz;
}
