// Host model check of the LDS exchange schedules of the FP64 transforms
// (cipher-pu_amd/csrc/ntt_f64.h) for every built launch shape.  Between two
// workgroup barriers the waves of a workgroup may drift arbitrarily, so an
// LDS slot touched there by two different waves, at least one of them
// writing, is a race.  The check replays each driver's sequence of LDS
// writes, reads and syncs -- the write/read slots from the headers' own
// layout and padding functions (ntt::index_of, ntt::pad, ntt::xshift) and
// the barriers where the headers' own predicates put them
// (ntt::wave_local, ntt::relaid_after_wave_local, nttf::y_relaid,
// nttf::y_relaid_inv) -- and reports every such slot.  Prints one line per
// failing (shape, schedule) and a summary; exit status 0 when none races.
#include <cstdio>
#include <vector>

#include "ntt_f64.h"

using ntt::Cfg;

static long failures = 0, checks = 0;

// Slot ownership inside the current barrier segment, per buffer word.
struct Lds {
    int words;
    std::vector<int> owner;       // -1 untouched, -2 several waves, else the wave
    std::vector<char> written;
    long races = 0;
    explicit Lds(int n) : words(n), owner(n, -1), written(n, 0) {}
    void barrier() {
        std::fill(owner.begin(), owner.end(), -1);
        std::fill(written.begin(), written.end(), 0);
    }
    void touch(int slot, int wave, bool write) {
        ++checks;
        int& o = owner[slot];
        if (o == -1) o = wave;
        else if (o != wave) {
            if (write || written[slot]) ++races;
            o = -2;
        }
        if (write) written[slot] = 1;
    }
};

// Every thread's slots of pass P's layout under padding K, at word offset base.
template <class C, int P, int K>
void access(Lds& m, int base, bool write) {
    constexpr int M = 1 << C::rsize(P), G = C::E >> C::rsize(P);
    for (int tid = 0; tid < C::T; ++tid)
        for (int g = 0; g < G; ++g)
            for (int e = 0; e < M; ++e) m.touch(base + ntt::pad<C, K>(ntt::index_of<C, P>(g, e, tid)), tid / 64, write);
}

template <bool W>
void sync(Lds& m) {
    if (!W) m.barrier();   // a wave-level sync orders nothing between waves
}

// nttf::fwd_rest2 / fwd_ntt2 (software-pipelined pair), rounds P .. NP-1.
template <class C, int P>
void fwd2(Lds& m) {
    if constexpr (P < C::NP) {
        constexpr bool W = ntt::wave_local<C, P>();
        constexpr int K = ntt::xshift<C>(P);
        const int bx = 0, by = C::LDS_WORDS;
        if constexpr (P == 1) m.barrier();
        else if constexpr (ntt::relaid_after_wave_local<C, P - 1, P>()) m.barrier();
        access<C, P - 1, K>(m, bx, true);
        sync<W>(m);
        access<C, P, K>(m, bx, false);
        if constexpr (nttf::y_relaid<C, P>()) m.barrier();
        access<C, P - 1, K>(m, by, true);
        sync<W>(m);
        access<C, P, K>(m, by, false);
        fwd2<C, P + 1>(m);
    }
}

// nttf::inv_rest2 / inv_ntt2, rounds P down to 0.
template <class C, int P>
void inv2(Lds& m) {
    if constexpr (P >= 0) {
        constexpr bool W = ntt::wave_local<C, P + 1>();
        constexpr int K = ntt::xshift<C>(P + 1);
        const int bx = 0, by = C::LDS_WORDS;
        if constexpr (P == C::NP - 2) m.barrier();
        else if constexpr (ntt::relaid_after_wave_local<C, P + 2, P + 1>()) m.barrier();
        access<C, P + 1, K>(m, bx, true);
        sync<W>(m);
        access<C, P, K>(m, bx, false);
        if constexpr (nttf::y_relaid_inv<C, P + 1>()) m.barrier();
        access<C, P + 1, K>(m, by, true);
        sync<W>(m);
        access<C, P, K>(m, by, false);
        inv2<C, P - 1>(m);
    }
}

// nttf::fwd_rest / inv_rest (one buffer, a sync after the writes and after the reads).
template <class C, int P>
void fwd1(Lds& m) {
    if constexpr (P < C::NP) {
        constexpr bool W = ntt::wave_local<C, P>();
        constexpr int K = ntt::xshift<C>(P);
        if constexpr (P == 1) m.barrier();
        else if constexpr (ntt::relaid_after_wave_local<C, P - 1, P>()) m.barrier();
        access<C, P - 1, K>(m, 0, true);
        sync<W>(m);
        access<C, P, K>(m, 0, false);
        sync<W>(m);
        fwd1<C, P + 1>(m);
    }
}

template <class C, int P>
void inv1(Lds& m) {
    if constexpr (P >= 0) {
        constexpr bool W = ntt::wave_local<C, P + 1>();
        constexpr int K = ntt::xshift<C>(P + 1);
        if constexpr (ntt::relaid_after_wave_local<C, P + 2, P + 1>()) m.barrier();
        access<C, P + 1, K>(m, 0, true);
        sync<W>(m);
        access<C, P, K>(m, 0, false);
        sync<W>(m);
        inv1<C, P - 1>(m);
    }
}

// nttf::fwd_rest_db / inv_rest_db (two alternating buffers, one sync each).
template <class C, int P>
void fwd_db(Lds& m) {
    if constexpr (P < C::NP) {
        constexpr bool W = ntt::wave_local<C, P>();
        if constexpr (P == 1) m.barrier();
        else if constexpr (P >= 3 && ntt::wave_local<C, P - 1>()) m.barrier();
        const int buf = (P & 1) ? C::LDS_WORDS : 0;
        access<C, P - 1, ntt::xshift<C>(P)>(m, buf, true);
        sync<W>(m);
        access<C, P, ntt::xshift<C>(P)>(m, buf, false);
        fwd_db<C, P + 1>(m);
    }
}

template <class C, int P>
void inv_db(Lds& m) {
    if constexpr (P >= 0) {
        constexpr bool W = ntt::wave_local<C, P + 1>();
        if constexpr (P == C::NP - 2) m.barrier();
        else if constexpr (P <= C::NP - 4 && ntt::wave_local<C, P + 2>()) m.barrier();
        const int buf = (P & 1) ? C::LDS_WORDS : 0;
        access<C, P + 1, ntt::xshift<C>(P + 1)>(m, buf, true);
        sync<W>(m);
        access<C, P, ntt::xshift<C>(P + 1)>(m, buf, false);
        inv_db<C, P - 1>(m);
    }
}

// The sequences the kernels run, transform after transform in one workgroup.
template <class C>
void shape(const char* name) {
    auto report = [&](const char* what, Lds& m) {
        if (m.races) {
            ++failures;
            std::printf("RACE %s %s: %ld slot accesses\n", name, what, m.races);
        }
    };
    {   // single transforms back to back (standalone transforms, key-switch digits without pairs)
        Lds m(C::LDS_WORDS);
        fwd1<C, 1>(m);
        fwd1<C, 1>(m);
        inv1<C, C::NP - 2>(m);
        fwd1<C, 1>(m);
        report("single", m);
    }
    if constexpr (ntt::pair_fits<C>()) {
        Lds m(2 * C::LDS_WORDS);
        // front: two forward pairs, then the double-buffered inverse;
        // key switch: double-buffered digit transforms, a barrier, the inverse pair
        fwd2<C, 1>(m);
        fwd2<C, 1>(m);
        inv_db<C, C::NP - 2>(m);
        fwd_db<C, 1>(m);
        fwd_db<C, 1>(m);
        m.barrier();
        inv2<C, C::NP - 2>(m);
        fwd2<C, 1>(m);
        report("pairs and double-buffered", m);
    }
}

int main() {
    shape<Cfg<12, 128>>("N=4096 T=128");
    shape<Cfg<12, 256>>("N=4096 T=256");
    shape<Cfg<12, 512>>("N=4096 T=512");
    shape<Cfg<13, 256>>("N=8192 T=256");
    shape<Cfg<13, 512>>("N=8192 T=512");
    shape<Cfg<13, 1024>>("N=8192 T=1024");
    shape<Cfg<14, 512>>("N=16384 T=512");
    shape<Cfg<14, 1024>>("N=16384 T=1024");
    std::printf("%ld checks, %ld failures\n", checks, failures);
    return failures ? 1 : 0;
}
