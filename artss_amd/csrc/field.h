// Field: a scalar grid function resident in HBM.
//
// Same public surface as the reference's Field (field/Field.h:17-147):
// public `data` pointer (here a device pointer), type/level/size getters,
// copy_data / set_value / += / *= as whole-array operations, and the pointer
// swap the V-cycle relies on.  Memory is zero-initialised (the reference
// leaves several fields uninitialised, Field.cpp:12-26).
#pragma once

#include <utility>

#include "domain.h"
#include "kernels.h"

namespace artss {

class Field {
  public:
    Field(Domain &d, int type, double val = 0.0, int level = 0)
        : m_domain(&d), m_type(type), m_level(level), m_size(d.level(level).size), m_owner(true) {
        data = d.alloc_field(level);
        if (val != 0.0) set_value(val);
    }
    // non-owning view of caller memory (the C ABI entry points use these)
    Field(Domain &d, int type, int level, double *external)
        : data(external), m_domain(&d), m_type(type), m_level(level), m_size(d.level(level).size), m_owner(false) {}
    ~Field() {
        if (m_owner) m_domain->free_field(data);
    }
    Field(const Field &) = delete;
    Field &operator=(const Field &) = delete;

    int get_type() const { return m_type; }
    int get_level() const { return m_level; }
    long long get_size() const { return m_size; }
    Domain &domain() const { return *m_domain; }

    void set_value(double v) const { k::fill(m_domain->stream(), data, v, m_size); }
    void copy_data(const Field &other) const { k::copy(m_domain->stream(), data, other.data, m_size); }
    static void swap(Field &a, Field &b) { std::swap(a.data, b.data); }
    Field &operator+=(const Field &rhs) {
        k::add(m_domain->stream(), data, rhs.data, m_size);
        return *this;
    }
    Field &operator*=(double x) {
        k::scale(m_domain->stream(), data, x, m_size);
        return *this;
    }

    double *data = nullptr;

  private:
    Domain *m_domain;
    int m_type;
    int m_level;
    long long m_size;
    bool m_owner;
};

}  // namespace artss
