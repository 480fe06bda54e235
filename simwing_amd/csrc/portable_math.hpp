// Portable double-precision atan2, sin, cos and acos.
//
// Every function here is built only from IEEE-754 operations that are
// correctly rounded on both the host CPU and gfx950 (+, -, *, /, sqrt and
// fused multiply-add), in a fixed order, with no table lookups other than the
// constant tables below. Compiled with floating-point contraction off, the
// host and the device therefore return identical bits for every input. The
// device kernels use these functions wherever the substep needs a
// transcendental, so a device run can be compared bit for bit with a CPU run
// that uses the same functions.
//
// Method (table-driven reduction in double-double arithmetic, after Tang,
// "Table-lookup algorithms for elementary functions", ACM TOMS 1990s, and
// Dekker/Knuth error-free transformations):
//   Quotients use one division and a Newton-corrected reciprocal.
//   atan(t), t in [0, 1]:  atan(t) = atan(c) + atan((t - c) / (1 + t c)),
//       c = k/64; the reduced argument |u| <= 2^-7 goes through the Maclaurin
//       series with the u^3 term in double-double.
//   sin/cos(x), |x| < 2^-9: the Maclaurin series in double (see sincos).
//   sin/cos(x): x = n pi/2 + r (three-part Cody-Waite constant, exact for
//       |n| < 2^20), r = j/64 + d, |d| <= 2^-7, and the angle-addition
//       formulas with tabulated sin(j/64), cos(j/64).
//   acos(x) = atan2(sqrt((1 - x)(1 + x)), x) with the radicand and the root
//       in double-double.
// The relative error before the final rounding is below 2^-75 for
// |x| < 2^20 pi/2 (sin, cos) and everywhere for atan2 and acos, so the
// results are correctly rounded except with probability below about 2^-20
// per call. Larger sin/cos arguments stay deterministic but lose accuracy.
#pragma once

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define SWX_PM __host__ __device__ inline
#else
#include <cmath>
#define SWX_PM inline
#endif
#define SWX_PM_TABLE inline constexpr

namespace swx {
namespace pm {

// BEGIN GENERATED by portable_math_tables.py
constexpr double kPiHi = 0x1.921fb54442d18p+1, kPiLo = 0x1.1a62633145c07p-53;
constexpr double kPio2Hi = 0x1.921fb54442d18p+0, kPio2Lo = 0x1.1a62633145c07p-54;
constexpr double kPio4 = 0x1.921fb54442d18p-1, k3Pio4 = 0x1.2d97c7f3321d2p+1;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
constexpr double kThirdHi = 0x1.5555555555555p-2, kThirdLo = 0x1.5555555555555p-56;
constexpr double kSixthHi = 0x1.5555555555555p-3, kSixthLo = 0x1.5555555555555p-57;
// pi/2 = kPio2Part1 + kPio2Part2 + kPio2Part3 + kPio2Part3Tail; parts 1 and 2 carry 33
// significant bits, so n * part is exact for |n| < 2^20.
constexpr double kPio2Part1 = 0x1.921fb54400000p+0, kPio2Part2 = 0x1.0b4611a600000p-34;
constexpr double kPio2Part3 = 0x1.3198a2e037073p-69, kPio2Part3Tail = 0x1.129024e088a68p-123;
// atan(k / 64), k = 0..64, as (hi, lo).
SWX_PM_TABLE double kAtanTable[65][2] = {
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
// sin(j / 64) and cos(j / 64), j = 0..51, as (sin hi, sin lo, cos hi, cos lo).
SWX_PM_TABLE double kSinCosTable[52][4] = {
    {0x0.0p+0, 0x0.0p+0, 0x1.0000000000000p+0, 0x0.0p+0},
    {0x1.fffaaaaeeeed5p-7, -0x1.2ab639a9f0776p-63, 0x1.fff000155549fp-1, 0x1.28a28a03a5ef3p-55},
    {0x1.ffeaaaeeee86fp-6, -0x1.cd406fb224ae2p-60, 0x1.ffc00155527d3p-1, -0x1.3b54492d89b5bp-55},
    {0x1.7fdc01032fba9p-5, -0x1.599bdf46e997ap-59, 0x1.ff7006bfdf99fp-1, -0x1.8b3b560648d5fp-56},
    {0x1.ffaaaeeed4edbp-5, -0x1.2d16d32684b69p-59, 0x1.ff0015549f4d3p-1, 0x1.328387b99426fp-55},
    {0x1.3facb12d1755bp-4, -0x1.921915299468bp-58, 0x1.fe7034129ef6fp-1, -0x1.cbf4337c96f97p-57},
    {0x1.7f701032550e4p-4, 0x1.afc2d1800501ap-60, 0x1.fdc06bf7e6b9bp-1, 0x1.31902b535f8dbp-55},
    {0x1.bf1b78568391dp-4, 0x1.e91841dea4cc8p-58, 0x1.fcf0c800e99b1p-1, 0x1.ea3d786d186acp-57},
    {0x1.feaaeee86ee36p-4, -0x1.afcb2bcc6f03bp-59, 0x1.fc015527d5bd3p-1, 0x1.b68f35094efb8p-55},
    {0x1.1f0d3d7afceafp-3, -0x1.6ef95099769a5p-57, 0x1.faf22263c4bd3p-1, -0x1.52ace133a2769p-58},
    {0x1.3eb312c5d66cbp-3, 0x1.47d666b66cb91p-57, 0x1.f9c340a7cc428p-1, 0x1.c5b6b063b7462p-55},
    {0x1.5e44fcfa126f3p-3, -0x1.6f443063f89b6p-57, 0x1.f874c2e1eecf6p-1, -0x1.c6514e1332b16p-55},
    {0x1.7dc102fbaf2b5p-3, 0x1.5ab50e23c97c3p-59, 0x1.f706bdf9ece1cp-1, -0x1.698c80c36dcb4p-55},
    {0x1.9d252d0cec312p-3, 0x1.9c43d80b1137dp-58, 0x1.f57948cff6797p-1, 0x1.e3a0d3e03b1d4p-57},
    {0x1.bc6f84edc6199p-3, 0x1.9c1a56a7b0cabp-57, 0x1.f3cc7c3b3d16ep-1, -0x1.21a3ad28a3494p-57},
    {0x1.db9e15fb5a5d0p-3, -0x1.32e20d6cc6fc2p-57, 0x1.f20073086649fp-1, 0x1.b940416c1984bp-56},
    {0x1.faaeed4f31577p-3, -0x1.15d88508e32b8p-57, 0x1.f01549f7deea1p-1, 0x1.d3c1e99e5cafdp-55},
    {0x1.0cd00cef36436p-2, -0x1.9fb0a0c93e2b4p-56, 0x1.ee0b1fbc0f11cp-1, -0x1.bfd2380bbc3b1p-59},
    {0x1.1c37d64c6b876p-2, 0x1.46076fe0dcff4p-56, 0x1.ebe214f76efa8p-1, -0x1.02f9f12ba543ep-55},
    {0x1.2b8ddc43eb49fp-2, 0x1.1553899f2d807p-57, 0x1.e99a4c3a7cd83p-1, -0x1.2264b1bc53ce8p-55},
    {0x1.3ad129769d3d8p-2, 0x1.03d550487839ap-63, 0x1.e733ea0193d40p-1, -0x1.6428b3546ce13p-55},
    {0x1.4a00c9b0f3d20p-2, 0x1.823ba6bb08eadp-56, 0x1.e4af14b2a449cp-1, -0x1.68ca02e8a6833p-55},
    {0x1.591bc9fa2f597p-2, 0x1.7c74bac3fe0cbp-57, 0x1.e20bf49acd6c1p-1, -0x1.660aec7ef636bp-58},
    {0x1.682138a38d7f7p-2, -0x1.d889202444aadp-56, 0x1.df4ab3ebd875ep-1, -0x1.e2d8a7e6736c4p-55},
    {0x1.7710255764214p-2, -0x1.6ead7314bb6cep-57, 0x1.dc6b7eb995912p-1, 0x1.4b364776dcd35p-58},
    {0x1.85e7a12826949p-2, 0x1.8a40e9b5face0p-56, 0x1.d96e82f71a9dcp-1, 0x1.ff61bd5d2039dp-55},
    {0x1.94a6be9f546c5p-2, -0x1.69ce13e683f58p-56, 0x1.d653f073e4040p-1, -0x1.76236434bec37p-55},
    {0x1.a34c91cc50ccap-2, -0x1.a310e3b50cecdp-58, 0x1.d31bf8d8d7c06p-1, 0x1.e60dd3089cbddp-56},
    {0x1.b1d8305321617p-2, -0x1.ae242cb99f519p-56, 0x1.cfc6cfa52ad9fp-1, 0x1.8b5b5508f2a0dp-55},
    {0x1.c048b17b140a3p-2, 0x1.19fe6757e9fa7p-57, 0x1.cc54aa2b2972ep-1, 0x1.4ee162ba83a98p-57},
    {0x1.ce9d2e3d4a51fp-2, -0x1.2fc8a12dae298p-57, 0x1.c8c5bf8ce1a84p-1, 0x1.ab3d1a1590123p-56},
    {0x1.dcd4c15329c9ap-2, 0x1.0d4c6e171fd9ap-56, 0x1.c51a48b8b175ep-1, -0x1.1bbb43b9aa880p-57},
    {0x1.eaee8744b05f0p-2, -0x1.789b43c9b027dp-58, 0x1.c1528065b7d50p-1, -0x1.892111312e828p-55},
    {0x1.f8e99e76abc97p-2, 0x1.9d950af2d00a3p-58, 0x1.bd6ea310294f5p-1, 0x1.31bbcc88c109dp-56},
    {0x1.0362939c69955p-1, -0x1.2d8cd78397b01p-55, 0x1.b96eeef58840ep-1, 0x1.45a3cc78fade0p-58},
    {0x1.0a4021e9e1001p-1, -0x1.6f643a13914f6p-55, 0x1.b553a410c104ep-1, 0x1.8ff7947027a15p-58},
    {0x1.110d0c4b69c3bp-1, 0x1.d918998809981p-55, 0x1.b11d04162a4c6p-1, 0x1.1dd561efbc0c2p-56},
    {0x1.17c8e5f2eedb0p-1, 0x1.35e57102e2488p-57, 0x1.accb526f69de5p-1, 0x1.8fb6a8dd6b6ccp-55},
    {0x1.1e7343236574cp-1, 0x1.22a3fa4f41d5ap-56, 0x1.a85ed4373e02dp-1, 0x1.9be06385ec792p-57},
    {0x1.250bb93788bbbp-1, 0x1.ea3d02457bccep-56, 0x1.a3d7d0352bdcfp-1, -0x1.68dbaeca19669p-55},
    {0x1.2b91dea88421ep-1, -0x1.fa371db216ab0p-55, 0x1.9f368ed912f85p-1, -0x1.1d200c5791606p-55},
    {0x1.32054b148bc4fp-1, 0x1.f6b42095a135bp-55, 0x1.9a7b5a36a6514p-1, 0x1.722cfcc9fa7a9p-55},
    {0x1.386597456282bp-1, -0x1.10fada93b07a8p-56, 0x1.95a67e00cb1fdp-1, -0x1.0befda21f862dp-55},
    {0x1.3eb25d36cd53ap-1, -0x1.be570e1570fc0p-58, 0x1.90b84784ddaf7p-1, -0x1.0feb10ab93b87p-56},
    {0x1.44eb381cf386bp-1, -0x1.3ed6c1e6a5505p-55, 0x1.8bb105a5dc900p-1, 0x1.863e03e9474c1p-55},
    {0x1.4b0fc46aab761p-1, 0x1.0da05738cc59cp-61, 0x1.869108d77a6c6p-1, 0x1.338ffe2bfe9ddp-56},
    {0x1.511f9fd7b351cp-1, -0x1.5c0e861c48831p-55, 0x1.8158a31916d5dp-1, -0x1.de8b90b8228dep-57},
    {0x1.571a6966d59b3p-1, 0x1.c843b4d0fb197p-58, 0x1.7c0827f09e54fp-1, -0x1.c73d6d72aee68p-57},
    {0x1.5cffc16bf8f0dp-1, 0x1.96cb370eb578ap-55, 0x1.769fec655211fp-1, -0x1.827d5cf8c68c5p-57},
    {0x1.62cf49921ac79p-1, -0x1.edd9855b6241ap-55, 0x1.712046fa77678p-1, 0x1.425b0a5029c81p-55},
    {0x1.6888a4e134b2fp-1, -0x1.6b7d37644d5e6p-55, 0x1.6b898fa9efb5dp-1, 0x1.15ac786ccf4b2p-56},
    {0x1.6e2b77c40bde1p-1, -0x1.0e729857fad53p-56, 0x1.65dc1fdeb8cbap-1, -0x1.97c1b47337c77p-58},
};
// END GENERATED

constexpr double kMaxFinite = 0x1.fffffffffffffp+1023;

struct dd {
    double hi, lo;
};

// Error-free transformations.
SWX_PM dd twoSum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
SWX_PM dd fastTwoSum(double a, double b) {  // |a| >= |b| or a == 0
    const double s = a + b;
    return {s, b - (s - a)};
}
SWX_PM dd twoProd(double a, double b) {
    const double p = a * b;
    return {p, fma(a, b, -p)};
}

SWX_PM dd add(dd a, dd b) {
    dd s = twoSum(a.hi, b.hi);
    const dd t = twoSum(a.lo, b.lo);
    s.lo += t.hi;
    s = fastTwoSum(s.hi, s.lo);
    s.lo += t.lo;
    return fastTwoSum(s.hi, s.lo);
}
SWX_PM dd neg(dd a) { return {-a.hi, -a.lo}; }
SWX_PM dd sub(dd a, dd b) { return add(a, neg(b)); }
SWX_PM dd addD(dd a, double b) {
    dd s = twoSum(a.hi, b);
    s.lo += a.lo;
    return fastTwoSum(s.hi, s.lo);
}
SWX_PM dd mul(dd a, dd b) {
    dd p = twoProd(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return fastTwoSum(p.hi, p.lo);
}
SWX_PM dd mulD(dd a, double b) {
    dd p = twoProd(a.hi, b);
    p.lo += a.lo * b;
    return fastTwoSum(p.hi, p.lo);
}
// 1 / b to about 2^-104 relative: one division and a Newton correction.
SWX_PM dd recip(dd b) {
    const double r = 1.0 / b.hi;
    const double e = fma(-b.hi, r, 1.0) - b.lo * r;
    return fastTwoSum(r, r * e);
}
// a / b to about 2^-103 relative.
SWX_PM dd quotient(dd a, dd b) { return mul(a, recip(b)); }
SWX_PM dd sqrtDD(dd a) {  // a >= 0
    if (a.hi <= 0.0) return {0.0, 0.0};
    const double s = sqrt(a.hi);
    const dd s2 = twoProd(s, s);
    const double e = ((a.hi - s2.hi) - s2.lo + a.lo) / (2.0 * s);
    return fastTwoSum(s, e);
}

// atan of t = (th, tl) in [0, 1] as a double-double. `table` holds
// kAtanTable's values (a device caller may pass a copy in LDS).
SWX_PM dd atanUnit(dd t, const double (*table)[2] = kAtanTable) {
    const int k = static_cast<int>(t.hi * 64.0 + 0.5);
    const double c = k * 0.015625;
    const dd num = twoSum(t.hi - c, t.lo);  // t.hi - c is exact (Sterbenz)
    dd den = twoProd(t.hi, c);
    den.lo += t.lo * c;
    den = addD(fastTwoSum(den.hi, den.lo), 1.0);
    const dd u = quotient(num, den);
    // atan(u) = u + u * (-u^2/3 + u^4/5 - ... + u^16/17), |u| <= 2^-7.
    const dd u2 = mul(u, u);
    const double s = u2.hi;
    const double tail =
        s * s *
        (1.0 / 5.0 -
         s * (1.0 / 7.0 - s * (1.0 / 9.0 - s * (1.0 / 11.0 - s * (1.0 / 13.0 - s * (1.0 / 15.0 - s * (1.0 / 17.0)))))));
    const dd w = addD(neg(mul(u2, dd{kThirdHi, kThirdLo})), tail);
    dd uw = twoProd(u.hi, w.hi);
    uw.lo += u.hi * w.lo + u.lo * w.hi;
    const dd series = add(u, fastTwoSum(uw.hi, uw.lo));
    return add(dd{table[k][0], table[k][1]}, series);
}

// atan2 of non-negative double-double operands, both finite and not both zero.
SWX_PM dd atan2Positive(dd y, dd x, const double (*table)[2] = kAtanTable) {
    const bool swap = y.hi > x.hi;
    const dd t = swap ? quotient(x, y) : quotient(y, x);
    dd r = atanUnit(t, table);
    if (swap) r = sub(dd{kPio2Hi, kPio2Lo}, r);
    return r;
}

SWX_PM double roundDD(dd a) { return a.hi + a.lo; }

// atan2(y, x) with the special values of C99 Annex F.
SWX_PM double atan2(double y, double x, const double (*table)[2] = kAtanTable) {
    if (x != x || y != y) return x + y;
    const double ax = fabs(x), ay = fabs(y);
    const bool negX = copysign(1.0, x) < 0.0;
    const bool infX = ax > kMaxFinite, infY = ay > kMaxFinite;
    double r;
    if (ay == 0.0) {
        r = negX ? kPiHi : 0.0;
    } else if (infX || infY) {
        if (infX && infY) r = negX ? k3Pio4 : kPio4;
        else if (infY) r = kPio2Hi;
        else r = negX ? kPiHi : 0.0;
    } else if (ax == 0.0) {
        r = kPio2Hi;
    } else {
        dd a = atan2Positive(dd{ay, 0.0}, dd{ax, 0.0}, table);
        if (negX) a = sub(dd{kPiHi, kPiLo}, a);
        r = roundDD(a);
    }
    return copysign(r, y);
}

// acos(x); NaN outside [-1, 1].
SWX_PM double acos(double x) {
    if (!(fabs(x) <= 1.0)) return (x - x) / (x - x);
    const dd oneMinus = twoSum(1.0, -x);
    const dd onePlus = twoSum(1.0, x);
    const dd root = sqrtDD(mul(oneMinus, onePlus));
    if (root.hi == 0.0) return x > 0.0 ? 0.0 : kPiHi;
    dd a = atan2Positive(root, dd{fabs(x), 0.0});
    if (x < 0.0) a = sub(dd{kPiHi, kPiLo}, a);
    return roundDD(a);
}

// sin and cos of x together.
SWX_PM void sincos(double x, double* sine, double* cosine) {
    if (!(fabs(x) <= kMaxFinite)) {  // NaN or infinite
        *sine = *cosine = (x - x) / (x - x);
        return;
    }
    if (fabs(x) < 0x1p-9) {
        // Small arguments (the rotation increments of the payload): the
        // Maclaurin series in double with one final rounding each,
        // sin x = x + x p(x^2) by a fused multiply-add and
        // cos x = 1 + (x^4 q(x^2) - x^2 / 2). The terms after the leading one
        // are below 2^-19 of it, so the error before the final rounding is
        // below 2^-72 (relative): correctly rounded except with probability
        // about 2^-19 per call. sin keeps the sign of zero.
        const double x2 = x * x;
        const double p = x2 * (-1.0 / 6.0 + x2 * (1.0 / 120.0 + x2 * (-1.0 / 5040.0 + x2 * (1.0 / 362880.0))));
        const double q =
            (x2 * x2) * (1.0 / 24.0 + x2 * (-1.0 / 720.0 + x2 * (1.0 / 40320.0 + x2 * (-1.0 / 3628800.0))));
        *sine = x == 0.0 ? x : fma(x, p, x);
        *cosine = 1.0 + (q - 0.5 * x2);
        return;
    }
    // x = n pi/2 + r. Exact reduction needs |n| < 2^20; beyond that the
    // result stays deterministic but loses accuracy.
    dd r{x, 0.0};
    int quadrant = 0;
    if (fabs(x) > kPio4) {
        const double n = rint(x * kTwoOverPi);
        const double r1 = x - n * kPio2Part1;
        r = twoSum(r1, -(n * kPio2Part2));
        const dd p3 = twoProd(n, kPio2Part3);
        r = sub(r, dd{p3.hi, p3.lo + n * kPio2Part3Tail});
        quadrant = static_cast<int>(n - 4.0 * floor(n * 0.25));
    }
    // r = j/64 + d, with j kept inside the table for out-of-range arguments.
    double jr = rint(r.hi * 64.0);
    jr = jr > 51.0 ? 51.0 : (jr < -51.0 ? -51.0 : jr);
    const int j = static_cast<int>(jr);
    const int aj = j < 0 ? -j : j;
    const dd d = twoSum(r.hi - j * 0.015625, r.lo);
    const double sgn = j < 0 ? -1.0 : 1.0;
    const dd sj{sgn * kSinCosTable[aj][0], sgn * kSinCosTable[aj][1]};
    const dd cj{kSinCosTable[aj][2], kSinCosTable[aj][3]};
    // sin d = d + d (-d^2/6 + d^4/5! - ...), cos d = 1 - d^2/2 + (d^4/4! - ...).
    const dd d2 = mul(d, d);
    const double s = d2.hi;
    const double sinTail =
        s * s * (1.0 / 120.0 - s * (1.0 / 5040.0 - s * (1.0 / 362880.0 - s * (1.0 / 39916800.0))));
    const dd ws = addD(neg(mul(d2, dd{kSixthHi, kSixthLo})), sinTail);
    dd dw = twoProd(d.hi, ws.hi);
    dw.lo += d.hi * ws.lo + d.lo * ws.hi;
    const dd sinD = add(d, fastTwoSum(dw.hi, dw.lo));
    const double cosTail =
        s * s * (1.0 / 24.0 - s * (1.0 / 720.0 - s * (1.0 / 40320.0 - s * (1.0 / 3628800.0 - s * (1.0 / 479001600.0)))));
    const dd cosD = addD(sub(dd{1.0, 0.0}, dd{0.5 * d2.hi, 0.5 * d2.lo}), cosTail);
    // j == 0 (|r| < 1/128): sin j/64 = 0 and cos j/64 = 1 make the
    // angle-addition formulas exact identities.
    const dd sinR = j == 0 ? sinD : add(mul(sj, cosD), mul(cj, sinD));
    const dd cosR = j == 0 ? cosD : sub(mul(cj, cosD), mul(sj, sinD));
    double sv, cv;
    switch (quadrant) {
    case 0: sv = roundDD(sinR); cv = roundDD(cosR); break;
    case 1: sv = roundDD(cosR); cv = -roundDD(sinR); break;
    case 2: sv = -roundDD(sinR); cv = -roundDD(cosR); break;
    default: sv = -roundDD(cosR); cv = roundDD(sinR); break;
    }
    if (x == 0.0) sv = x;  // keeps the sign of zero
    *sine = sv;
    *cosine = cv;
}

SWX_PM double sin(double x) {
    double s, c;
    sincos(x, &s, &c);
    return s;
}
SWX_PM double cos(double x) {
    double s, c;
    sincos(x, &s, &c);
    return c;
}

}  // namespace pm
}  // namespace swx
