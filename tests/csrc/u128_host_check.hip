// Host build of the device arithmetic (benchntt_amd/csrc/u128.hpp) so that the
// CPU test suite can compare it with the oracle before any GPU run.  Every
// function of u128.hpp is __host__ __device__ and uses no device builtins.
#include <stddef.h>
#include <stdint.h>

#include "../../benchntt_amd/csrc/u128.hpp"

using bntt::U128;

extern "C" {

// op: 0 addmod_ref, 1 submod_ref, 2 mulmod_ref, 3 mulmod_pm, 4 mulmod<true>, 5 mulmod<false>,
//     6 mulmod_r31 (b converted to digits first), 7 montmul_r31 (mu = -q^-1 mod 2^124),
//     8 addmod_canon, 9 submod_canon, 10 montmul_r31_lazy (mu = -q^-1 mod 2^124),
//     11 one lazy Montgomery butterfly of x = a, y = b with w = c (out[2i], out[2i+1]),
//     12 montmul_r31_lazy<SPARSE> (q's 31-bit digits 1 and 2 zero),
//     13 addmod_pm, 14 submod_pm (m of the form 2^124 - c: only its limb 0 is read)
void chk_op(int op, U128 *out, const U128 *a, const U128 *b, size_t n, const U128 *m, const U128 *mu, uint32_t c)
{
    bntt::Mod md;
    md.m = *m;
    md.mu = *mu;
    for (int i = 0; i < 4; i++) md.notm.v[i] = ~m->v[i];
    md.c = c;
    md.pm = c != 0;
    for (size_t i = 0; i < n; i++) {
        switch (op) {
        case 0: out[i] = bntt::addmod(a[i], b[i], md); break;
        case 1: out[i] = bntt::submod(a[i], b[i], md); break;
        case 2: out[i] = bntt::mulmod_ref(a[i], b[i], md.m, md.mu); break;
        case 3: out[i] = bntt::mulmod_pm(a[i], b[i], c); break;
        case 4: out[i] = bntt::mulmod<true>(a[i], b[i], md); break;
        case 5: out[i] = bntt::mulmod<false>(a[i], b[i], md); break;
        case 6: out[i] = bntt::mulmod_r31(a[i], bntt::to_r31(b[i]), c); break;
        case 7: {
            bntt::MontQ mq;
            mq.q = bntt::to_r31(*m);
            mq.qinv = bntt::to_r31(*mu);
            mq.ql = *m;
            out[i] = bntt::montmul_r31(a[i], bntt::to_r31(b[i]), mq);
            break;
        }
        case 8: out[i] = bntt::addmod_canon(a[i], b[i], *m); break;
        case 13: out[i] = bntt::addmod_pm(a[i], b[i], m->v[0]); break;
        case 14: out[i] = bntt::submod_pm(a[i], b[i], m->v[0]); break;
        case 10:
        case 11:
        case 12: {
            bntt::MontQ mq;
            mq.q = bntt::to_r31(*m);
            mq.qinv = bntt::to_r31(*mu);
            mq.ql = *m;
            mq.ql2 = bntt::add128l(*m, *m);
            if (op == 10) {
                out[i] = bntt::montmul_r31_lazy(a[i], bntt::to_r31(b[i]), mq);
            } else if (op == 12) {
                out[i] = bntt::montmul_r31_lazy<true>(a[i], bntt::to_r31(b[i]), mq);
            } else {
                // the butterfly of reg_network_mont on (x = a, y = b), twiddle digit 0 = c
                bntt::R31 w{};
                w.d[0] = c;
                const U128 x = bntt::sub_if_ge(a[i], mq.ql2);
                const U128 p = bntt::montmul_r31_lazy(b[i], w, mq);
                out[2 * i] = bntt::add128l(x, p);
                out[2 * i + 1] = bntt::add128l(x, bntt::sub128l(mq.ql2, p));
            }
            break;
        }
        default: out[i] = bntt::submod_canon(a[i], b[i], *m); break;
        }
    }
}
}
