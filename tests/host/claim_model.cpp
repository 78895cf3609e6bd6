// CPU model of the Tier-1 wave schedulers (test infrastructure).
//
// Waves are state machines stepped by an explicit scheduler, so every
// interleaving of their atomic steps can be explored.  The hardware
// dispatcher is modelled as in-order dispatch into a bounded number of
// resident slots: wave i becomes resident only after waves 0..i-1 did and
// while fewer than `resident` waves are alive; a wave frees its slot when
// it exits.  Waves are placed on SIMDs round-robin.
//
// Protocols:
//   pick   the shipped hand-out (t1_pick_block in kernels_t1.hip): a wave
//          counts itself on its SIMD; the first wave on a SIMD takes the next
//          position from the front of the cost order, any later one the next
//          from the back.  The first wave decides after a wait bounded by the
//          clock (heavy unless the launch's rule is on and its SIMD has three
//          arrivals by then); the explorer lets the decision fall anywhere
//          between the other waves' steps and takes both outcomes, so every
//          way the wait can end and either setting of the rule is covered.
//          One block per wave, n waves.  No wave waits on another.
//   claim  persistent waves: each claims positions from one counter until
//          none are left.  No wave waits.
//   mark   the exclusive-SIMD form: the first wave on a SIMD claims a
//          position and publishes whether it is heavy; a later wave on that
//          SIMD waits for the publication, leaves if heavy, else claims like
//          `claim`.  The first wave publishes only after its claim succeeds.
//   split  `pick` with the counter increments split into a read and a
//          separate write (not atomic): a planted mistake, which the
//          explorer must report as a position coded twice.
//
// For each protocol the explorer runs every interleaving (depth-first over
// which runnable wave steps next, each distinct state once) for small sizes
// and reports: every position coded exactly once on every complete run, and
// whether some interleaving reaches a state where no wave can step while
// waves remain (a hang).
//
// Usage: claim_model <protocol> <waves> <blocks> <simds> <resident>
// Prints "ok <states>" or "hang <trace>" or "wrong <trace>" (the trace: d =
// a dispatch, a digit = a step of that wave); exit status 0 for ok, 1 for a
// hang, 2 for a coverage error.  `pick` needs as many waves as blocks.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

namespace {

enum Proto { PICK, CLAIM, MARK, SPLIT };

struct Wave {
    int pc = 0;         // program counter of the wave's state machine
    int simd = 0;
    bool first = false;
    int pos = -1;       // position being coded
    int seen = 0;       // value read by a split increment
    bool done = false;  // exited
    bool started = false;
};

struct State {
    std::vector<Wave> w;
    std::vector<int> simd_count;  // arrivals per SIMD (the per-SIMD counter)
    std::vector<int> mark;        // -1 unpublished, 0 light, 1 heavy
    int front = 0, back = 0, next = 0;
    std::vector<int> coded;       // times each position was coded
    int alive = 0, dispatched = 0;
};

std::string key(const State& s) {
    std::string k;
    for (const Wave& v : s.w) {
        k.push_back((char)v.pc);
        k.push_back((char)v.simd);
        k.push_back((char)(v.first | (v.done << 1) | (v.started << 2)));
        k.push_back((char)(v.pos + 1));
        k.push_back((char)v.seen);
    }
    for (int c : s.simd_count) k.push_back((char)c);
    for (int c : s.mark) k.push_back((char)(c + 1));
    for (int c : s.coded) k.push_back((char)c);
    k.push_back((char)s.front);
    k.push_back((char)s.back);
    k.push_back((char)s.next);
    k.push_back((char)s.alive);
    k.push_back((char)s.dispatched);
    return k;
}

struct Model {
    Proto proto;
    int waves, blocks, simds, resident, heavy;
    long runs = 0;
    std::unordered_set<std::string> seen;
    std::string bad;
    int status = 0;  // 0 ok, 1 hang, 2 wrong

    bool can_dispatch(const State& s) const { return s.dispatched < waves && s.alive < resident; }

    // One atomic step of wave i; returns false if the wave is blocked.
    bool step(State& s, int i) const {
        Wave& v = s.w[i];
        switch (proto) {
        case PICK:
            if (v.pc == 0) {  // atomicAdd on the SIMD's arrival counter
                v.first = s.simd_count[v.simd]++ == 0;
                v.pc = v.first ? 1 : 2;
            } else if (v.pc == 1) {  // first wave, after its bounded wait
                return false;             // heavy or light: both explored (explore)
            } else if (v.pc == 2) {  // atomicAdd on the front or back counter
                v.pos = v.first ? s.front++ : blocks - 1 - s.back++;
                v.pc = 3;
            } else {  // code the block, exit
                if (v.pos >= 0 && v.pos < blocks) s.coded[v.pos]++;
                v.done = true;
            }
            return true;
        case SPLIT:
            if (v.pc == 0) {
                v.first = s.simd_count[v.simd]++ == 0;
                v.pc = 1;
            } else if (v.pc == 1) {  // read the counter ...
                v.seen = v.first ? s.front : s.back;
                v.pc = 2;
            } else if (v.pc == 2) {  // ... and write it back incremented
                (v.first ? s.front : s.back) = v.seen + 1;
                v.pos = v.first ? v.seen : blocks - 1 - v.seen;
                v.pc = 3;
            } else {
                if (v.pos >= 0 && v.pos < blocks) s.coded[v.pos]++;
                v.done = true;
            }
            return true;
        case CLAIM:
            if (v.pc == 0) {
                v.pos = s.next++;
                v.pc = v.pos < blocks ? 1 : 2;
            } else if (v.pc == 1) {
                s.coded[v.pos]++;
                v.pc = 0;
            } else {
                v.done = true;
            }
            return true;
        case MARK:
            if (v.pc == 0) {
                v.first = s.simd_count[v.simd]++ == 0;
                v.pc = v.first ? 1 : 4;
            } else if (v.pc == 1) {  // first wave: claim, then publish
                v.pos = s.next++;
                v.pc = v.pos < blocks ? 2 : 9;  // nothing left: exits unpublished
            } else if (v.pc == 2) {
                s.mark[v.simd] = v.pos < heavy ? 1 : 0;
                v.pc = 3;
            } else if (v.pc == 3) {  // code, then keep claiming
                s.coded[v.pos]++;
                v.pc = 5;
            } else if (v.pc == 4) {  // later wave: wait for the publication
                if (s.mark[v.simd] < 0) return false;
                v.pc = s.mark[v.simd] == 1 ? 9 : 5;
            } else if (v.pc == 5) {
                v.pos = s.next++;
                v.pc = v.pos < blocks ? 6 : 9;
            } else if (v.pc == 6) {
                s.coded[v.pos]++;
                v.pc = 5;
            } else {
                v.done = true;
            }
            return true;
        }
        return true;
    }

    void finish(const State& s, const std::string& trace) {
        runs++;
        for (int b = 0; b < blocks; b++)
            if (s.coded[b] != 1 && status == 0) {
                status = 2;
                bad = trace + " position " + std::to_string(b) + " coded " + std::to_string(s.coded[b]) + "x";
            }
    }

    void explore(State& s, std::string& trace) {
        if (status) return;
        if (!seen.insert(key(s)).second) return;
        bool any = false;
        // dispatch the next wave (in order) while a slot is free
        if (can_dispatch(s)) {
            State t = s;
            Wave& v = t.w[t.dispatched];
            v.started = true;
            v.simd = t.dispatched % simds;
            t.dispatched++;
            t.alive++;
            trace.push_back('d');
            explore(t, trace);
            trace.pop_back();
            any = true;
        }
        for (int i = 0; i < s.dispatched && !status; i++) {
            if (s.w[i].done) continue;
            if (proto == PICK && s.w[i].pc == 1) {
                // the first wave's decision: heavy (front) or light (back),
                // whatever its SIMD's count and the launch's rule say
                for (int heavy = 0; heavy < 2 && !status; heavy++) {
                    State t = s;
                    t.w[i].first = heavy != 0;
                    t.w[i].pc = 2;
                    trace.push_back((char)('0' + i));
                    explore(t, trace);
                    trace.pop_back();
                }
                any = true;
                continue;
            }
            State t = s;
            if (!step(t, i)) continue;
            if (t.w[i].done) t.alive--;
            trace.push_back((char)('0' + i));
            explore(t, trace);
            trace.pop_back();
            any = true;
        }
        if (!any && !status) {
            if (s.dispatched == waves && s.alive == 0) {
                finish(s, trace);
            } else {
                status = 1;
                bad = trace;
            }
        }
    }
};

}  // namespace

int main(int argc, char** argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: claim_model pick|claim|mark|split waves blocks simds resident\n");
        return 3;
    }
    Model m;
    const std::string p = argv[1];
    m.proto = p == "pick" ? PICK : p == "claim" ? CLAIM : p == "split" ? SPLIT : MARK;
    m.waves = atoi(argv[2]);
    m.blocks = atoi(argv[3]);
    m.simds = atoi(argv[4]);
    m.resident = atoi(argv[5]);
    m.heavy = 1;  // position 0 is heavy
    if ((m.proto == PICK || m.proto == SPLIT) && m.waves != m.blocks) {
        fprintf(stderr, "pick launches one wave per block\n");
        return 3;
    }
    State s;
    s.w.resize(m.waves);
    s.simd_count.assign(m.simds, 0);
    s.mark.assign(m.simds, -1);
    s.coded.assign(m.blocks, 0);
    std::string trace;
    m.explore(s, trace);
    if (m.status == 0) {
        printf("ok %zu\n", m.seen.size());
        return 0;
    }
    printf("%s %s\n", m.status == 1 ? "hang" : "wrong", m.bad.c_str());
    return m.status;
}
