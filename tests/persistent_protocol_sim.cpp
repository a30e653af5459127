// Host model of the scheduling protocol of ffbp_persistent_kernel
// (bp_project_amd/csrc/bpmi_ffbp.hip) -- test infrastructure, never linked
// into the library.
//
// G workgroups run on R resident slots (a workgroup starts when a slot is
// free, as the hardware dispatcher starts a launch's workgroups); each runs
// the kernel's loop:
//   claim PW_CHUNK consecutive items with an atomic ticket (nothing more once
//   the abort flag is set); for each item: wait for its dependencies (the
//   children's completion counts of a merge tile, the root's of an
//   evaluation block) -- remembering the last pair seen complete, flushing
//   the pending count before the first re-check, giving up after the wall
//   clock limit by setting the abort flag; run the item (a random delay);
//   add the completion to a pending count kept per counter and flushed when
//   the counter changes, before a wait and at the end of the chunk.
// Every access to the shared counters is a sequentially consistent atomic:
// the model checks the protocol's logic (ordering, pending counts, claims,
// residency), not the memory model of the device.
//
// Input (stdin): nodes items jobs, then per node "left right tiles", per item
// "kind idx", per job "root blocks".  Arguments: G R seed max_wait_ms
// max_delay_us [noflush].  "noflush" plants a mistake in the protocol: no
// flush of the pending count before a wait.  Output: one line
// "executed E aborted A bad B exited X twice T short S".
// Exit status 0 when every item ran once, nothing aborted, every dependency
// was complete when its consumer started and every counter reached its
// target.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

namespace {

constexpr int CHUNK = 16;
enum { LEAF = 0, MERGE = 1, EVAL = 2 };

struct Model {
    int nodes = 0, nitems = 0, njobs = 0;
    std::vector<int> left, right, tiles, kind, idx, job_root, job_blocks;
    std::vector<std::atomic<int>> done;  // node counters, then job counters
    std::vector<std::atomic<int>> ran;   // per item: times executed
    std::atomic<int> ticket{0}, abort{0}, bad{0}, exited{0};
    double max_wait_ms = 1000.0;
    int max_delay_us = 20;
    bool no_flush_before_wait = false;
};

struct Wg {
    int pend_idx = -1, pend_cnt = 0, ready0 = -2, ready1 = -2;
};

void flush(Model& m, Wg& w) {
    if (w.pend_cnt <= 0) return;
    m.done[w.pend_idx].fetch_add(w.pend_cnt);
    w.pend_cnt = 0;
}

int target(const Model& m, int counter) {
    return counter < m.nodes ? m.tiles[counter] : m.job_blocks[counter - m.nodes];
}

bool wait_item(Model& m, Wg& w, int i) {
    int n0 = -1, n1 = -1;
    if (m.kind[i] == MERGE) {
        n0 = m.left[m.idx[i]];
        n1 = m.right[m.idx[i]];
    } else if (m.kind[i] == EVAL) {
        n0 = m.job_root[m.idx[i]];
    }
    if (n0 < 0 || (n0 == w.ready0 && n1 == w.ready1)) return true;
    const int need0 = m.tiles[n0], need1 = n1 >= 0 ? m.tiles[n1] : 0;
    const auto t0 = std::chrono::steady_clock::now();
    for (bool first = true;; first = false) {
        if (m.done[n0].load() >= need0 && (n1 < 0 || m.done[n1].load() >= need1)) {
            w.ready0 = n0;
            w.ready1 = n1;
            return true;
        }
        if (first && !m.no_flush_before_wait) flush(m, w);
        if (m.abort.load()) return false;
        const double ms = std::chrono::duration<double, std::milli>(
                              std::chrono::steady_clock::now() - t0).count();
        if (ms > m.max_wait_ms) {
            m.abort.store(1);
            return false;
        }
        std::this_thread::yield();
    }
}

bool deps_complete(const Model& m, int i) {
    if (m.kind[i] == MERGE) {
        const int l = m.left[m.idx[i]], r = m.right[m.idx[i]];
        return m.done[l].load() >= m.tiles[l] && m.done[r].load() >= m.tiles[r];
    }
    if (m.kind[i] == EVAL) {
        const int root = m.job_root[m.idx[i]];
        return m.done[root].load() >= m.tiles[root];
    }
    return true;
}

void run_workgroup(Model& m, std::mt19937& rng) {
    Wg w;
    std::uniform_int_distribution<int> delay(0, m.max_delay_us);
    for (;;) {
        const int first = m.abort.load() ? m.nitems : m.ticket.fetch_add(CHUNK);
        if (first >= m.nitems) break;
        const int last = std::min(first + CHUNK, m.nitems);
        for (int i = first; i < last; ++i) {
            const bool go = wait_item(m, w, i);
            if (go) {
                if (!deps_complete(m, i)) m.bad.fetch_add(1);
                const int us = delay(rng);
                if (us) std::this_thread::sleep_for(std::chrono::microseconds(us));
                m.ran[i].fetch_add(1);
                const int counter = m.kind[i] == EVAL ? m.nodes + m.idx[i] : m.idx[i];
                if (counter != w.pend_idx) {
                    flush(m, w);
                    w.pend_idx = counter;
                }
                w.pend_cnt += 1;
            }
        }
        flush(m, w);
    }
    m.exited.fetch_add(1);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 6 && argc != 7) {
        std::fprintf(stderr, "usage: %s G R seed max_wait_ms max_delay_us [noflush] < worklist\n",
                      argv[0]);
        return 2;
    }
    const int G = std::atoi(argv[1]), R = std::atoi(argv[2]);
    const unsigned seed = (unsigned)std::atoi(argv[3]);
    Model m;
    m.max_wait_ms = std::atof(argv[4]);
    m.max_delay_us = std::atoi(argv[5]);
    m.no_flush_before_wait = argc == 7;
    if (std::scanf("%d %d %d", &m.nodes, &m.nitems, &m.njobs) != 3) return 2;
    m.left.resize(m.nodes);
    m.right.resize(m.nodes);
    m.tiles.resize(m.nodes);
    for (int n = 0; n < m.nodes; ++n)
        if (std::scanf("%d %d %d", &m.left[n], &m.right[n], &m.tiles[n]) != 3) return 2;
    m.kind.resize(m.nitems);
    m.idx.resize(m.nitems);
    for (int i = 0; i < m.nitems; ++i)
        if (std::scanf("%d %d", &m.kind[i], &m.idx[i]) != 2) return 2;
    m.job_root.resize(m.njobs);
    m.job_blocks.resize(m.njobs);
    for (int j = 0; j < m.njobs; ++j)
        if (std::scanf("%d %d", &m.job_root[j], &m.job_blocks[j]) != 2) return 2;
    m.done = std::vector<std::atomic<int>>(m.nodes + m.njobs);
    m.ran = std::vector<std::atomic<int>>(m.nitems);

    // R slots; a slot runs workgroups one after another until G have started
    std::atomic<int> started{0};
    std::vector<std::thread> slots;
    for (int s = 0; s < std::min(R, G); ++s)
        slots.emplace_back([&, s] {
            std::mt19937 rng(seed * 7919u + (unsigned)s);
            while (started.fetch_add(1) < G) run_workgroup(m, rng);
        });
    for (auto& t : slots) t.join();

    int executed = 0, twice = 0;
    for (int i = 0; i < m.nitems; ++i) {
        executed += m.ran[i].load() > 0;
        twice += m.ran[i].load() > 1;
    }
    int short_counts = 0;
    for (int c = 0; c < m.nodes + m.njobs; ++c) short_counts += m.done[c].load() != target(m, c);
    std::printf("executed %d aborted %d bad %d exited %d twice %d short %d\n", executed,
                m.abort.load(), m.bad.load(), m.exited.load(), twice, short_counts);
    const bool ok = executed == m.nitems && !m.abort.load() && !m.bad.load() &&
                    m.exited.load() == G && !twice && !short_counts;
    return ok ? 0 : 1;
}
