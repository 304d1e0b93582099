#include "dyn_smagorinsky.h"

#include <cmath>

namespace artss {

DynamicSmagorinsky::DynamicSmagorinsky(Domain &d) : m_domain(d) {
    for (int q = 0; q < COUNT; q++) m_work.push_back(std::make_unique<Field>(d, U));
}

void DynamicSmagorinsky::explicit_filtering(Field &out, const Field &in, bool sync) {
    k::DynArrays o{};
    k::DynConstArrays i{};
    o.p[0] = out.data;
    i.p[0] = in.data;
    k::dyn_filter(m_domain.stream(), m_domain.level(0), o, i, 1);
    if (sync) ARTSS_HIP(hipStreamSynchronize(m_domain.stream()));
}

// DynamicSmagorinsky.cpp:172-393
void DynamicSmagorinsky::calc_turbulent_viscosity(Field &ev, const Field &u, const Field &v, const Field &wv,
                                                  bool sync) {
    Domain &d = m_domain;
    const LevelDesc &L = d.level(0);
    hipStream_t s = d.stream();
    const double dx = L.h[0], dy = L.h[1], dz = L.h[2];
    const double delta_s = std::cbrt(dx * dy * dz);
    const double alpha = 2.0;
    const double c2 = 2.0 * delta_s * delta_s;  // the left part of :330-335, in the reference's order
    const double a2 = alpha * alpha;
    // velocity filter, :199-201
    {
        k::DynArrays o{};
        k::DynConstArrays i{};
        o.p[0] = w(UF), o.p[1] = w(VF), o.p[2] = w(WF);
        i.p[0] = u.data, i.p[1] = v.data, i.p[2] = wv.data;
        k::dyn_filter(s, L, o, i, 3);
    }
    // products, :212-228
    {
        k::DynArrays p{};
        for (int c = 0; c < 6; c++) p.p[c] = w(PROD + c);
        k::dyn_products(s, L, w(UF), w(VF), w(WF), p);
    }
    // filtered products, :235-240: uu vv ww uv vw (uw_f is left as it is)
    {
        k::DynArrays o{};
        k::DynConstArrays i{};
        const int which[5] = {0, 1, 2, 3, 5};
        for (int c = 0; c < 5; c++) {
            o.p[c] = w(FPROD + which[c]);
            i.p[c] = w(PROD + which[c]);
        }
        k::dyn_filter(s, L, o, i, 5);
    }
    // strain, modulus and products, :256-282
    {
        k::DynArrays o{};
        for (int c = 0; c < 12; c++) o.p[c] = w(S + c);
        k::dyn_strain(s, L, w(UF), w(VF), w(WF), o, w(SBAR));
    }
    // filtered strain and products, :289-303
    {
        k::DynArrays o{};
        k::DynConstArrays i{};
        for (int c = 0; c < 12; c++) {
            o.p[c] = w(SF + c);
            i.p[c] = w(S + c);
        }
        k::dyn_filter(s, L, o, i, 12);
    }
    // coefficient, :308-350
    {
        k::DynConstArrays f{}, p{}, sf{};
        for (int c = 0; c < 6; c++) {
            f.p[c] = w(FPROD + c);
            p.p[c] = w(PROD + c);
        }
        for (int c = 0; c < 12; c++) sf.p[c] = w(SF + c);
        k::dyn_coef(s, L, c2, a2, f, p, sf, w(CS));
    }
    // in-place average, eddy viscosity and clamp, :358-392
    k::dyn_average(s, L, delta_s, d.settings().nu, w(CS), w(SBAR), ev.data);
    if (sync) ARTSS_HIP(hipStreamSynchronize(s));
}

}  // namespace artss
