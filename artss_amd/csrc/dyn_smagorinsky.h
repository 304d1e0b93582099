// turbulence/DynamicSmagorinsky.cpp:17-427: the dynamic Smagorinsky-Lilly
// eddy viscosity behind ITurbulence.
//
// Where the reference leaves values undefined this build defines them: its
// work fields are allocated without values and never written outside the
// inner cells, and uw_f is never written at all (its filter call filters ww
// twice, :235-240); here all of those read zero, which is what a fresh
// allocation of the reference reads.  The coefficient average (:372-384)
// runs in place in index order in the serial build; the build reproduces
// that order with one launch per wavefront i + 2 j + 4 k = t.
#pragma once

#include <memory>
#include <vector>

#include "operators.h"

namespace artss {

class DynamicSmagorinsky : public ITurbulence {
  public:
    explicit DynamicSmagorinsky(Domain &d);
    void calc_turbulent_viscosity(Field &ev, const Field &in_u, const Field &in_v, const Field &in_w,
                                  bool sync) override;
    // :405-427 on the inner fluid cells of `out`
    void explicit_filtering(Field &out, const Field &in, bool sync) override;

  private:
    Domain &m_domain;
    // u_f v_f w_f; products uu vv ww uv uw vw; filtered products (uw_f stays
    // zero); S and S_bar*S; their filtered copies; S_bar; Cs
    enum { UF, VF, WF, PROD, FPROD = PROD + 6, S = FPROD + 6, SF = S + 12, SBAR = SF + 12, CS, COUNT };
    std::vector<std::unique_ptr<Field>> m_work;
    double *w(int q) { return m_work[q]->data; }
};

}  // namespace artss
