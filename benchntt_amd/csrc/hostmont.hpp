// hostmont.hpp -- host-side arithmetic mod an odd q < 2^122 for building
// tables: the device's Montgomery product (R = 2^124, montmul_r31) compiled
// for the host, a Miller-Rabin check and the 2-power roots of unity.
#pragma once

#include <cstring>

#include "u128.hpp"

namespace bntt {

typedef unsigned __int128 u128h;

inline U128 u128_to_limbs(u128h v) { return make_u128((uint64_t)v, (uint64_t)(v >> 64)); }
inline u128h limbs_to_u128(const U128 &a) { return ((u128h)hi64(a) << 64) | lo64(a); }
inline uint32_t brv_host(uint32_t i, int bits) { return bits ? (__builtin_bitreverse32(i) >> (32 - bits)) : 0; }

struct HostMont {
    MontQ mq;
    U128 r1;  // R mod q: Montgomery form of 1
    R31 r2;   // R^2 mod q in digits
    u128h q = 0;

    // false unless q is odd, 3 <= q < 2^122
    bool init(u128h qv)
    {
        q = qv;
        if (!(q & 1) || (q >> 122) != 0 || q < 3) return false;
        // q^-1 mod 2^128 by Newton iteration, then -q^-1 mod 2^124
        u128h x = q;
        for (int i = 0; i < 7; i++) x *= (u128h)2 - q * x;
        const u128h qinv = ((~x + 1) & (((u128h)1 << 124) - 1));
        mq.q = to_r31(u128_to_limbs(q));
        mq.qinv = to_r31(u128_to_limbs(qinv));
        mq.ql = u128_to_limbs(q);
        mq.ql2 = u128_to_limbs(q << 1);
        // R mod q and R^2 mod q by doubling
        u128h v = 1;
        for (int i = 0; i < 248; i++) {
            v <<= 1;
            if (v >= q) v -= q;
            if (i == 123) r1 = u128_to_limbs(v);
        }
        r2 = to_r31(u128_to_limbs(v));
        return true;
    }
    U128 to_m(u128h x) const { return montmul_r31(u128_to_limbs(x % q), r2, mq); }
    U128 mul(const U128 &a, const U128 &b) const { return montmul_r31(a, to_r31(b), mq); }
    u128h from_m(const U128 &a) const
    {
        R31 one{};
        one.d[0] = 1;
        return limbs_to_u128(montmul_r31(a, one, mq));
    }
    U128 pow(U128 base, u128h e) const
    {
        U128 r = r1;
        while (e) {
            if (e & 1) r = mul(r, base);
            base = mul(base, base);
            e >>= 1;
        }
        return r;
    }
    bool eq(const U128 &a, const U128 &b) const { return std::memcmp(&a, &b, sizeof a) == 0; }

    // Miller-Rabin with the 16 prime bases 2..53.
    bool is_probable_prime() const
    {
        static const unsigned bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53};
        for (unsigned p : bases)
            if (q % p == 0) return q == p;
        u128h d = q - 1;
        int s = 0;
        while (!(d & 1)) {
            d >>= 1;
            s++;
        }
        const U128 one = r1, minus_one = to_m(q - 1);
        for (unsigned a : bases) {
            U128 x = pow(to_m(a), d);
            if (eq(x, one) || eq(x, minus_one)) continue;
            bool composite = true;
            for (int i = 1; i < s && composite; i++) {
                x = mul(x, x);
                if (eq(x, minus_one)) composite = false;
            }
            if (composite) return false;
        }
        return true;
    }

    // 2^v = largest power of two dividing q - 1; w (Montgomery form) a root
    // of unity of order exactly 2^v: g^((q-1)/2^v) for the least quadratic
    // non-residue g.  q must be prime.  False if no non-residue below 1000.
    bool two_adic_root(U128 *w, int *v) const
    {
        int e = 0;
        while ((((q - 1) >> e) & 1) == 0) e++;
        const U128 minus_one = to_m(q - 1);
        for (u128h g = 2; g < 1000; g++) {
            const U128 gm = to_m(g);
            if (eq(pow(gm, (q - 1) >> 1), minus_one)) {
                *w = pow(gm, (q - 1) >> e);
                *v = e;
                return true;
            }
        }
        return false;
    }
};

}  // namespace bntt
