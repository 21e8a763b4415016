// ne_atan2.hpp -- atan2 rounded to nearest, in double-double arithmetic, for
// the hinge limit's angle (src/world.cpp:300, std::atan2).
//
// The device library's atan2 differs from the host C library's in the last
// bit or two for about a quarter of the hinge limit's arguments.  This one
// carries about 104 bits to the final rounding, so it returns the correctly
// rounded atan2: the host library's result wherever that is correctly
// rounded.  The same code on the host and on the device gives the same bits
// (IEEE + - * / and fma only).
//
// Method: z = min(|y|, |x|) / max(|y|, |x|) as a double-double; atan(z) =
// atan(c) + atan((z - c) / (1 + z c)) with c = k/64 the nearest 64th, the
// second term by its odd series to t^17 (|t| <= 1/128); then the octant and
// quadrant by pi/2 - r and pi - r.  Constants (atan(k/64), 1/(2m+1), pi,
// pi/2 as double-doubles) come from ne_atan2_tables.py.
#pragma once

#include <math.h>

#ifndef NE_HD
#define NE_HD inline
#endif

namespace nea {
namespace atan2cr {

struct DD {
    double h, l;
};

NE_HD DD two_sum(double a, double b) {
    const double s = a + b, bb = s - a;
    return DD{s, (a - (s - bb)) + (b - bb)};
}
NE_HD DD fast_two_sum(double a, double b) {   // |a| >= |b|
    const double s = a + b;
    return DD{s, b - (s - a)};
}
NE_HD DD two_prod(double a, double b) {
    const double p = a * b;
    return DD{p, fma(a, b, -p)};
}
NE_HD DD add(DD x, DD y) {
    DD s = two_sum(x.h, y.h), t = two_sum(x.l, y.l);
    s.l += t.h;
    s = fast_two_sum(s.h, s.l);
    s.l += t.l;
    return fast_two_sum(s.h, s.l);
}
NE_HD DD neg(DD x) { return DD{-x.h, -x.l}; }
NE_HD DD mul(DD x, DD y) {
    DD p = two_prod(x.h, y.h);
    p.l += x.h * y.l + x.l * y.h;
    return fast_two_sum(p.h, p.l);
}
NE_HD DD mul_d(DD x, double c) {
    DD p = two_prod(x.h, c);
    p.l += x.l * c;
    return fast_two_sum(p.h, p.l);
}
NE_HD DD div(DD x, DD y) {
    const double q1 = x.h / y.h;
    DD r = add(x, neg(mul_d(y, q1)));
    const double q2 = r.h / y.h;
    r = add(r, neg(mul_d(y, q2)));
    const double q3 = r.h / y.h;
    return add(fast_two_sum(q1, q2), DD{q3, 0.0});
}

// atan(k/64), k = 0..64
constexpr double ATAN_TAB[65][2] = {
    {0x0.0p+0, 0x0.0p+0},
    {0x1.fff555bbb729bp-7, -0x1.220c39d4dff50p-61},
    {0x1.ffd55bba97625p-6, -0x1.5ec431444912cp-60},
    {0x1.7fb818430da2ap-5, -0x1.86ef8f794f105p-63},
    {0x1.ff55bb72cfdeap-5, -0x1.c934d86d23f1dp-60},
    {0x1.3f59f0e7c559dp-4, 0x1.ac4ce285df847p-58},
    {0x1.7ee182602f10fp-4, -0x1.cfb654c0c3d98p-58},
    {0x1.be39ebe6f07c3p-4, 0x1.f7b8f29a05987p-58},
    {0x1.fd5ba9aac2f6ep-4, -0x1.cd37686760c17p-59},
    {0x1.1e1fafb043727p-3, -0x1.b485914dacf8cp-59},
    {0x1.3d6eee8c6626cp-3, 0x1.61a3b0ce9281bp-57},
    {0x1.5c9811e3ec26ap-3, -0x1.054ab2c010f3dp-58},
    {0x1.7b97b4bce5b02p-3, 0x1.347b0b4f881cap-58},
    {0x1.9a6a8e96c8626p-3, 0x1.cf601e7b4348ep-59},
    {0x1.b90d7529260a2p-3, 0x1.17b10d2e0e5abp-61},
    {0x1.d77d5df205736p-3, 0x1.c648d1534597ep-57},
    {0x1.f5b75f92c80ddp-3, 0x1.8ab6e3cf7afbdp-57},
    {0x1.09dc597d86362p-2, 0x1.62e47390cb865p-56},
    {0x1.18bf5a30bf178p-2, 0x1.30ca4748b1bf9p-57},
    {0x1.278372057ef46p-2, -0x1.077cdd36dfc81p-56},
    {0x1.362773707ebccp-2, -0x1.963a544b672d8p-57},
    {0x1.44aa436c2af0ap-2, -0x1.5d5e43c55b3bap-56},
    {0x1.530ad9951cd4ap-2, -0x1.2566480884082p-57},
    {0x1.614840309cfe2p-2, -0x1.a725715711f00p-56},
    {0x1.6f61941e4def1p-2, -0x1.c63aae6f6e918p-56},
    {0x1.7d5604b63b3f7p-2, 0x1.69c885c2b249ap-56},
    {0x1.8b24d394a1b25p-2, 0x1.b6d0ba3748fa8p-56},
    {0x1.98cd5454d6b18p-2, 0x1.9e6c988fd0a77p-56},
    {0x1.a64eec3cc23fdp-2, -0x1.24dec1b50b7ffp-56},
    {0x1.b3a911da65c6cp-2, 0x1.ae187b1ca5040p-56},
    {0x1.c0db4c94ec9f0p-2, -0x1.cc1ce70934c34p-56},
    {0x1.cde53432c1351p-2, -0x1.a2cfa4418f1adp-56},
    {0x1.dac670561bb4fp-2, 0x1.a2b7f222f65e2p-56},
    {0x1.e77eb7f175a34p-2, 0x1.0e53dc1bf3435p-56},
    {0x1.f40dd0b541418p-2, -0x1.a3992dc382a23p-57},
    {0x1.0039c73c1a40cp-1, -0x1.b32c949c9d593p-55},
    {0x1.0657e94db30d0p-1, -0x1.d5b495f6349e6p-56},
    {0x1.0c6145b5b43dap-1, 0x1.974fa13b5404fp-58},
    {0x1.1255d9bfbd2a9p-1, -0x1.2bdaee1c0ee35p-58},
    {0x1.1835a88be7c13p-1, 0x1.c621cec00c301p-55},
    {0x1.1e00babdefeb4p-1, -0x1.928df287a668fp-58},
    {0x1.23b71e2cc9e6ap-1, 0x1.c421c9f38224ep-57},
    {0x1.2958e59308e31p-1, -0x1.09e73b0c6c087p-56},
    {0x1.2ee628406cbcap-1, 0x1.c5d5e9ff0cf8dp-55},
    {0x1.345f01cce37bbp-1, 0x1.1021137c71102p-55},
    {0x1.39c391cd4171ap-1, -0x1.2304331d8bf46p-55},
    {0x1.3f13fb89e96f4p-1, 0x1.ecf8b492644f0p-56},
    {0x1.445065b795b56p-1, -0x1.f76d0163f79c8p-56},
    {0x1.4978fa3269ee1p-1, 0x1.2419a87f2a458p-56},
    {0x1.4e8de5bb6ec04p-1, 0x1.4a33dbeb3796cp-55},
    {0x1.538f57b89061fp-1, -0x1.1bb74abda520cp-55},
    {0x1.587d81f732fbbp-1, -0x1.5e5c9d8c5a950p-56},
    {0x1.5d58987169b18p-1, 0x1.0028e4bc5e7cap-57},
    {0x1.6220d115d7b8ep-1, -0x1.2b785350ee8c1p-57},
    {0x1.66d663923e087p-1, -0x1.6ea6febe8bbbap-56},
    {0x1.6b798920b3d99p-1, -0x1.a80386188c50ep-55},
    {0x1.700a7c5784634p-1, -0x1.8c34d25aadef6p-56},
    {0x1.748978fba8e0fp-1, 0x1.7b2a6165884a1p-59},
    {0x1.78f6bbd5d315ep-1, 0x1.406a089803740p-55},
    {0x1.7d528289fa093p-1, 0x1.560821e2f3aa9p-55},
    {0x1.819d0b7158a4dp-1, -0x1.bf76229d3b917p-56},
    {0x1.85d69576cc2c5p-1, 0x1.6b66e7fc8b8c3p-57},
    {0x1.89ff5ff57f1f8p-1, -0x1.55b9a5e177a1bp-55},
    {0x1.8e17aa99cc05ep-1, -0x1.ec182ab042f61p-56},
    {0x1.921fb54442d18p-1, 0x1.1a62633145c07p-55},
};
// 1/(2m+1), m = 1..10
constexpr double INV_ODD[10][2] = {
    {0x1.5555555555555p-2, 0x1.5555555555555p-56},
    {0x1.999999999999ap-3, -0x1.999999999999ap-57},
    {0x1.2492492492492p-3, 0x1.2492492492492p-57},
    {0x1.c71c71c71c71cp-4, 0x1.c71c71c71c71cp-58},
    {0x1.745d1745d1746p-4, -0x1.745d1745d1746p-59},
    {0x1.3b13b13b13b14p-4, -0x1.3b13b13b13b14p-58},
    {0x1.1111111111111p-4, 0x1.1111111111111p-60},
    {0x1.e1e1e1e1e1e1ep-5, 0x1.e1e1e1e1e1e1ep-61},
    {0x1.af286bca1af28p-5, 0x1.af286bca1af28p-59},
    {0x1.8618618618618p-5, 0x1.8618618618618p-59},
};

// atan(z) for a double-double z in [0, 1]
NE_HD DD atan01(DD z, const double (*tab)[2], const double (*inv)[2]) {
    int k = (int)(z.h * 64.0 + 0.5);
    k = k < 0 ? 0 : (k > 64 ? 64 : k);
    const double c = k * (1.0 / 64.0);
    const DD num = add(z, DD{-c, 0.0});
    const DD den = add(DD{1.0, 0.0}, mul_d(z, c));
    const DD t = div(num, den);
    const DD s = mul(t, t);
    DD acc = DD{inv[7][0], inv[7][1]};   // 1/17, then Horner down to 1
    for (int m = 7; m >= 1; --m) {
        const DD cm = (m % 2) ? DD{-inv[m - 1][0], -inv[m - 1][1]} : DD{inv[m - 1][0], inv[m - 1][1]};
        acc = add(mul(acc, s), cm);
    }
    acc = add(mul(acc, s), DD{1.0, 0.0});
    return add(DD{tab[k][0], tab[k][1]}, mul(t, acc));
}

NE_HD double atan2_rn(double y, double x, const double (*tab)[2], const double (*inv)[2]) {
    const DD PI = DD{0x1.921fb54442d18p+1, 0x1.1a62633145c07p-53};
    const DD PI_2 = DD{0x1.921fb54442d18p+0, 0x1.1a62633145c07p-54};
    if (isnan(x) || isnan(y)) return x + y;
    const bool yneg = signbit(y), xneg = signbit(x);
    if (y == 0.0) return xneg ? (yneg ? -PI.h : PI.h) : y;   // atan2(+-0, -0 or x < 0) = +-pi, else +-0
    if (x == 0.0) return yneg ? -PI_2.h : PI_2.h;
    if (isinf(x) || isinf(y)) {
        double r;
        if (isinf(x) && isinf(y)) r = xneg ? 3.0 * (PI.h * 0.25) : PI.h * 0.25;   // 3pi/4 rounds from 3 * fl(pi/4)
        else if (isinf(y)) r = PI_2.h;
        else r = xneg ? PI.h : 0.0;
        return yneg ? -r : r;
    }
    const double ax = fabs(x), ay = fabs(y);
    const bool swap = ay > ax;
    const double num = swap ? ax : ay, den = swap ? ay : ax;
    const double zh = num / den;
    const DD z = DD{zh, fma(-zh, den, num) / den};
    DD r = atan01(fast_two_sum(z.h, z.l), tab, inv);
    if (swap) r = add(PI_2, neg(r));
    if (xneg) r = add(PI, neg(r));
    return yneg ? -r.h : r.h;
}

}  // namespace atan2cr

// The hinge limit's atan2, on the host or the device.
NE_HD double atan2_rn(double y, double x) { return atan2cr::atan2_rn(y, x, atan2cr::ATAN_TAB, atan2cr::INV_ODD); }

}  // namespace nea
