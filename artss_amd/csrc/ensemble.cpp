#include "ensemble.h"

#include <stdexcept>

namespace artss {

Ensemble::Ensemble(int threads) {
    if (threads < 1) throw std::invalid_argument("an ensemble needs at least one thread");
    for (int t = 1; t < threads; t++) m_pool.emplace_back([this] { worker(); });
}

Ensemble::~Ensemble() {
    {
        std::lock_guard<std::mutex> g(m_lock);
        m_stop = true;
    }
    m_wake.notify_all();
    for (std::thread &t : m_pool) t.join();
}

void Ensemble::add(Solver *member) {
    if (member == nullptr) throw std::invalid_argument("null ensemble member");
    for (Solver *m : m_members)
        if (m == member) throw std::invalid_argument("a solver can be an ensemble member only once");
    m_members.push_back(member);
}

// takes members until none is left; the calling thread works as well
void Ensemble::run_members() {
    while (true) {
        size_t i;
        long n;
        {
            std::lock_guard<std::mutex> g(m_lock);
            if (m_next >= m_members.size()) return;
            i = m_next++;
            n = m_step;
        }
        try {
            Solver &s = *m_members[i];
            s.step(static_cast<double>(n) * s.domain().settings().dt);
            // a member's step is complete on the device when step() returns
            ARTSS_HIP(hipStreamSynchronize(s.domain().stream()));
        } catch (...) {
            std::lock_guard<std::mutex> g(m_lock);
            if (!m_error) m_error = std::current_exception();
        }
    }
}

void Ensemble::worker() {
    long seen = 0;
    while (true) {
        {
            std::unique_lock<std::mutex> g(m_lock);
            m_wake.wait(g, [&] { return m_stop || m_generation != seen; });
            if (m_stop) return;
            seen = m_generation;
            m_running++;
        }
        run_members();
        {
            std::lock_guard<std::mutex> g(m_lock);
            m_running--;
        }
        m_done.notify_all();
    }
}

void Ensemble::step(long n) {
    {
        std::lock_guard<std::mutex> g(m_lock);
        m_step = n;
        m_next = 0;
        m_error = nullptr;
        m_generation++;
    }
    m_wake.notify_all();
    run_members();
    std::exception_ptr error;
    {
        std::unique_lock<std::mutex> g(m_lock);
        // every member has been taken; wait for the workers still inside one
        m_done.wait(g, [&] { return m_running == 0; });
        error = m_error;
    }
    if (error) std::rethrow_exception(error);
}

}  // namespace artss
