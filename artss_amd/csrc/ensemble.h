// Ensembles of independent solvers of uneven size (prognosis and data
// assimilation run many perturbed copies of one scenario).  It has no
// counterpart in the reference, which advances one domain per process.
//
// Every member keeps its own HIP stream, fields and multigrid hierarchy; a
// pool of host threads issues the members' steps concurrently, so kernels of
// different members overlap on the GPU and the host round trips of one
// member's residual checks hide behind the others' work.  Members are not
// owned and must outlive the ensemble's use of them.
#pragma once

#include <condition_variable>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

#include "solver.h"

namespace artss {

class Ensemble {
  public:
    explicit Ensemble(int threads);
    ~Ensemble();
    Ensemble(const Ensemble &) = delete;
    Ensemble &operator=(const Ensemble &) = delete;

    void add(Solver *member);
    size_t size() const { return m_members.size(); }
    Solver &member(size_t i) { return *m_members[i]; }

    // every member advances one step: member m runs Solver::step(n * dt_m)
    // with its own time step dt_m.  Members run concurrently on `threads`
    // host threads (threads = 1: one after another on the calling thread).
    // The first exception a member throws is rethrown after all are done.
    // Every member's stream is synchronised before step() returns.
    void step(long n);

  private:
    void worker();
    void run_members();

    std::vector<Solver *> m_members;
    std::vector<std::thread> m_pool;
    std::mutex m_lock;
    std::condition_variable m_wake, m_done;
    long m_generation = 0;  // bumped once per step
    long m_step = 0;
    size_t m_next = 0;      // next member to take
    size_t m_running = 0;   // workers inside the current step
    bool m_stop = false;
    std::exception_ptr m_error;
};

}  // namespace artss
