// exp and sincos in FP64, bit-equal to what the reference's Linux build (gcc, x86-64, glibc >= 2.28)
// calls in its Gaussian model (lib/fit/GaussianModel.cpp:19-52): std::exp of the Gaussian, and
// std::cos / std::sin of one angle, which gcc merges into one sincos call (the reference compiled
// here imports sincos and neither sin nor cos).
//
// exp: Tang's table-driven exponential with N = 128 as published in ARM's optimized-routines
// (math/exp.c), the one glibc ships: x = k ln2/N + r with |r| <= ln2/(2N), exp(x) = 2^(k/N) *
// exp(r), exp(r) - 1 by a degree-5 polynomial, 2^(i/N) = H_i * (1 + t_i) from a 128-entry table.
// glibc runs its FMA build of exp on CPUs with FMA: the reduction, the polynomial and the final
// scaling are fused multiply-adds, the subnormal path is not, and so here.
// Table: for i in [0, 128), H_i is the double nearest 2^(i/128) and t_i the double nearest
// 2^(i/128)/H_i - 1; entry 2i holds the bits of t_i, entry 2i+1 the bits of H_i minus i << 45
// (the shifted k adds them back).
//
// sincos: the IBM Accurate Mathematical Library method that glibc ships (sysdeps/ieee754/dbl-64
// s_sin.c, s_sincos.c): |x| < 2^-27 trivially; |x| < 0.855469 directly; |x| < 2.426265 through
// pi/2 - |x| (pi/2 as a double-double); |x| < 105414350 after a reduction by multiples of pi/2 split
// over four constants. The reduced argument a (+ tail da) is split at the nearest X_i = i/128, sin
// and cos of the rest by short odd / even polynomials and combined with sin(X_i), cos(X_i) from a
// double-double table (a Taylor-like polynomial for sin of |a| < 0.126 instead). glibc has no FMA
// build of sincos: every operation rounds separately, and so here. Table: for i in [0, 110),
// entries 4i..4i+3 are the double nearest sin(i/128), the double nearest the rest, and the same two
// for cos(i/128). Beyond 105414350 (no angle a fit produces) the device's own sin / cos are used.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace pt {

__device__ __constant__ static const uint64_t kExpTable[256] = {
    0x0000000000000000ull, 0x3ff0000000000000ull, 0x3c9b3b4f1a88bf6eull, 0x3feff63da9fb3335ull,
    0xbc7160139cd8dc5dull, 0x3fefec9a3e778061ull, 0xbc905e7a108766d1ull, 0x3fefe315e86e7f85ull,
    0x3c8cd2523567f613ull, 0x3fefd9b0d3158574ull, 0xbc8bce8023f98efaull, 0x3fefd06b29ddf6deull,
    0x3c60f74e61e6c861ull, 0x3fefc74518759bc8ull, 0x3c90a3e45b33d399ull, 0x3fefbe3ecac6f383ull,
    0x3c979aa65d837b6dull, 0x3fefb5586cf9890full, 0x3c8eb51a92fdeffcull, 0x3fefac922b7247f7ull,
    0x3c3ebe3d702f9cd1ull, 0x3fefa3ec32d3d1a2ull, 0xbc6a033489906e0bull, 0x3fef9b66affed31bull,
    0xbc9556522a2fbd0eull, 0x3fef9301d0125b51ull, 0xbc5080ef8c4eea55ull, 0x3fef8abdc06c31ccull,
    0xbc91c923b9d5f416ull, 0x3fef829aaea92de0ull, 0x3c80d3e3e95c55afull, 0x3fef7a98c8a58e51ull,
    0xbc801b15eaa59348ull, 0x3fef72b83c7d517bull, 0xbc8f1ff055de323dull, 0x3fef6af9388c8deaull,
    0x3c8b898c3f1353bfull, 0x3fef635beb6fcb75ull, 0xbc96d99c7611eb26ull, 0x3fef5be084045cd4ull,
    0x3c9aecf73e3a2f60ull, 0x3fef54873168b9aaull, 0xbc8fe782cb86389dull, 0x3fef4d5022fcd91dull,
    0x3c8a6f4144a6c38dull, 0x3fef463b88628cd6ull, 0x3c807a05b0e4047dull, 0x3fef3f49917ddc96ull,
    0x3c968efde3a8a894ull, 0x3fef387a6e756238ull, 0x3c875e18f274487dull, 0x3fef31ce4fb2a63full,
    0x3c80472b981fe7f2ull, 0x3fef2b4565e27cddull, 0xbc96b87b3f71085eull, 0x3fef24dfe1f56381ull,
    0x3c82f7e16d09ab31ull, 0x3fef1e9df51fdee1ull, 0xbc3d219b1a6fbffaull, 0x3fef187fd0dad990ull,
    0x3c8b3782720c0ab4ull, 0x3fef1285a6e4030bull, 0x3c6e149289cecb8full, 0x3fef0cafa93e2f56ull,
    0x3c834d754db0abb6ull, 0x3fef06fe0a31b715ull, 0x3c864201e2ac744cull, 0x3fef0170fc4cd831ull,
    0x3c8fdd395dd3f84aull, 0x3feefc08b26416ffull, 0xbc86a3803b8e5b04ull, 0x3feef6c55f929ff1ull,
    0xbc924aedcc4b5068ull, 0x3feef1a7373aa9cbull, 0xbc9907f81b512d8eull, 0x3feeecae6d05d866ull,
    0xbc71d1e83e9436d2ull, 0x3feee7db34e59ff7ull, 0xbc991919b3ce1b15ull, 0x3feee32dc313a8e5ull,
    0x3c859f48a72a4c6dull, 0x3feedea64c123422ull, 0xbc9312607a28698aull, 0x3feeda4504ac801cull,
    0xbc58a78f4817895bull, 0x3feed60a21f72e2aull, 0xbc7c2c9b67499a1bull, 0x3feed1f5d950a897ull,
    0x3c4363ed60c2ac11ull, 0x3feece086061892dull, 0x3c9666093b0664efull, 0x3feeca41ed1d0057ull,
    0x3c6ecce1daa10379ull, 0x3feec6a2b5c13cd0ull, 0x3c93ff8e3f0f1230ull, 0x3feec32af0d7d3deull,
    0x3c7690cebb7aafb0ull, 0x3feebfdad5362a27ull, 0x3c931dbdeb54e077ull, 0x3feebcb299fddd0dull,
    0xbc8f94340071a38eull, 0x3feeb9b2769d2ca7ull, 0xbc87deccdc93a349ull, 0x3feeb6daa2cf6642ull,
    0xbc78dec6bd0f385full, 0x3feeb42b569d4f82ull, 0xbc861246ec7b5cf6ull, 0x3feeb1a4ca5d920full,
    0x3c93350518fdd78eull, 0x3feeaf4736b527daull, 0x3c7b98b72f8a9b05ull, 0x3feead12d497c7fdull,
    0x3c9063e1e21c5409ull, 0x3feeab07dd485429ull, 0x3c34c7855019c6eaull, 0x3feea9268a5946b7ull,
    0x3c9432e62b64c035ull, 0x3feea76f15ad2148ull, 0xbc8ce44a6199769full, 0x3feea5e1b976dc09ull,
    0xbc8c33c53bef4da8ull, 0x3feea47eb03a5585ull, 0xbc845378892be9aeull, 0x3feea34634ccc320ull,
    0xbc93cedd78565858ull, 0x3feea23882552225ull, 0x3c5710aa807e1964ull, 0x3feea155d44ca973ull,
    0xbc93b3efbf5e2228ull, 0x3feea09e667f3bcdull, 0xbc6a12ad8734b982ull, 0x3feea012750bdabfull,
    0xbc6367efb86da9eeull, 0x3fee9fb23c651a2full, 0xbc80dc3d54e08851ull, 0x3fee9f7df9519484ull,
    0xbc781f647e5a3ecfull, 0x3fee9f75e8ec5f74ull, 0xbc86ee4ac08b7db0ull, 0x3fee9f9a48a58174ull,
    0xbc8619321e55e68aull, 0x3fee9feb564267c9ull, 0x3c909ccb5e09d4d3ull, 0x3feea0694fde5d3full,
    0xbc7b32dcb94da51dull, 0x3feea11473eb0187ull, 0x3c94ecfd5467c06bull, 0x3feea1ed0130c132ull,
    0x3c65ebe1abd66c55ull, 0x3feea2f336cf4e62ull, 0xbc88a1c52fb3cf42ull, 0x3feea427543e1a12ull,
    0xbc9369b6f13b3734ull, 0x3feea589994cce13ull, 0xbc805e843a19ff1eull, 0x3feea71a4623c7adull,
    0xbc94d450d872576eull, 0x3feea8d99b4492edull, 0x3c90ad675b0e8a00ull, 0x3feeaac7d98a6699ull,
    0x3c8db72fc1f0eab4ull, 0x3feeace5422aa0dbull, 0xbc65b6609cc5e7ffull, 0x3feeaf3216b5448cull,
    0x3c7bf68359f35f44ull, 0x3feeb1ae99157736ull, 0xbc93091fa71e3d83ull, 0x3feeb45b0b91ffc6ull,
    0xbc5da9b88b6c1e29ull, 0x3feeb737b0cdc5e5ull, 0xbc6c23f97c90b959ull, 0x3feeba44cbc8520full,
    0xbc92434322f4f9aaull, 0x3feebd829fde4e50ull, 0xbc85ca6cd7668e4bull, 0x3feec0f170ca07baull,
    0x3c71affc2b91ce27ull, 0x3feec49182a3f090ull, 0x3c6dd235e10a73bbull, 0x3feec86319e32323ull,
    0xbc87c50422622263ull, 0x3feecc667b5de565ull, 0x3c8b1c86e3e231d5ull, 0x3feed09bec4a2d33ull,
    0xbc91bbd1d3bcbb15ull, 0x3feed503b23e255dull, 0x3c90cc319cee31d2ull, 0x3feed99e1330b358ull,
    0x3c8469846e735ab3ull, 0x3feede6b5579fdbfull, 0xbc82dfcd978e9db4ull, 0x3feee36bbfd3f37aull,
    0x3c8c1a7792cb3387ull, 0x3feee89f995ad3adull, 0xbc907b8f4ad1d9faull, 0x3feeee07298db666ull,
    0xbc55c3d956dcaebaull, 0x3feef3a2b84f15fbull, 0xbc90a40e3da6f640ull, 0x3feef9728de5593aull,
    0xbc68d6f438ad9334ull, 0x3feeff76f2fb5e47ull, 0xbc91eee26b588a35ull, 0x3fef05b030a1064aull,
    0x3c74ffd70a5fddcdull, 0x3fef0c1e904bc1d2ull, 0xbc91bdfbfa9298acull, 0x3fef12c25bd71e09ull,
    0x3c736eae30af0cb3ull, 0x3fef199bdd85529cull, 0x3c8ee3325c9ffd94ull, 0x3fef20ab5fffd07aull,
    0x3c84e08fd10959acull, 0x3fef27f12e57d14bull, 0x3c63cdaf384e1a67ull, 0x3fef2f6d9406e7b5ull,
    0x3c676b2c6c921968ull, 0x3fef3720dcef9069ull, 0xbc808a1883ccb5d2ull, 0x3fef3f0b555dc3faull,
    0xbc8fad5d3ffffa6full, 0x3fef472d4a07897cull, 0xbc900dae3875a949ull, 0x3fef4f87080d89f2ull,
    0x3c74a385a63d07a7ull, 0x3fef5818dcfba487ull, 0xbc82919e2040220full, 0x3fef60e316c98398ull,
    0x3c8e5a50d5c192acull, 0x3fef69e603db3285ull, 0x3c843a59ac016b4bull, 0x3fef7321f301b460ull,
    0xbc82d52107b43e1full, 0x3fef7c97337b9b5full, 0xbc892ab93b470dc9ull, 0x3fef864614f5a129ull,
    0x3c74b604603a88d3ull, 0x3fef902ee78b3ff6ull, 0x3c83c5ec519d7271ull, 0x3fef9a51fbc74c83ull,
    0xbc8ff7128fd391f0ull, 0x3fefa4afa2a490daull, 0xbc8dae98e223747dull, 0x3fefaf482d8e67f1ull,
    0x3c8ec3bc41aa2008ull, 0x3fefba1bee615a27ull, 0x3c842b94c3a9eb32ull, 0x3fefc52b376bba97ull,
    0x3c8a64a931d185eeull, 0x3fefd0765b6e4540ull, 0xbc8e37bae43be3edull, 0x3fefdbfdad9cbe14ull,
    0x3c77893b4d91cd9dull, 0x3fefe7c1819e90d8ull, 0x3c5305c14160cc89ull, 0x3feff3c22b8f71f1ull,
};

__device__ inline double libm_exp(double x)
{
#pragma clang fp contract(off)
    constexpr double kInvLn2N = 0x1.71547652b82fep7, kShift = 0x1.8p52;
    constexpr double kNegLn2HiN = -0x1.62e42fefa0000p-8, kNegLn2LoN = -0x1.cf79abc9e3b3ap-47;
    constexpr double kC2 = 0x1.ffffffffffdbdp-2, kC3 = 0x1.555555555543cp-3, kC4 = 0x1.55555cf172b91p-5,
                     kC5 = 0x1.1111167a4d017p-7;
    const uint64_t ux = (uint64_t)__double_as_longlong(x);
    uint32_t abstop = (uint32_t)(ux >> 52) & 0x7ffu;
    // |x| < 2^-54, |x| >= 512 or not finite: top12(2^-54) = 0x3c9, top12(512) = 0x408.
    if (abstop - 0x3c9u >= 0x408u - 0x3c9u) {
        if (abstop - 0x3c9u >= 0x80000000u) return 1.0 + x;   // tiny |x|
        if (abstop >= 0x409u) {                                // |x| >= 1024, inf, nan
            if (ux == 0xfff0000000000000ull) return 0.0;
            if (abstop >= 0x7ffu) return 1.0 + x;
            return (ux >> 63) ? 0.0 : __longlong_as_double(0x7ff0000000000000ll);
        }
        abstop = 0;   // 512 <= |x| < 1024: the scaled path below
    }
    double kd = __builtin_fma(kInvLn2N, x, kShift);
    const uint64_t ki = (uint64_t)__double_as_longlong(kd);
    kd -= kShift;
    const double r = __builtin_fma(kd, kNegLn2LoN, __builtin_fma(kd, kNegLn2HiN, x));
    const uint32_t idx = 2u * (uint32_t)(ki & 127u);
    const double tail = __longlong_as_double((long long)kExpTable[idx]);
    uint64_t sbits = kExpTable[idx + 1] + (ki << 45);
    const double r2 = r * r;
    const double tmp =
        __builtin_fma(r2 * r2, __builtin_fma(r, kC5, kC4), __builtin_fma(r2, __builtin_fma(r, kC3, kC2), tail + r));
    if (abstop == 0) {
        if ((ki & 0x80000000u) == 0) {   // k > 0: the scale's exponent may overflow by up to 460
            sbits -= 1009ull << 52;
            const double scale = __longlong_as_double((long long)sbits);
            return 0x1p1009 * __builtin_fma(scale, tmp, scale);
        }
        sbits += 1022ull << 52;           // k < 0: round once in the subnormal range
        const double scale = __longlong_as_double((long long)sbits);
        double y = scale + scale * tmp;
        if (y < 1.0) {
            double lo = scale - y + scale * tmp;
            const double hi = 1.0 + y;
            lo = 1.0 - hi + y + lo;
            y = (hi + lo) - 1.0;
            if (y == 0.0) y = 0.0;
        }
        return 0x1p-1022 * y;
    }
    const double scale = __longlong_as_double((long long)sbits);
    return __builtin_fma(scale, tmp, scale);
}

__device__ __constant__ static const uint64_t kSinCosTable[440] = {
    0x0000000000000000ull, 0x0000000000000000ull, 0x3ff0000000000000ull, 0x0000000000000000ull,
    0x3f7fffeaaaaeeeefull, 0xbc1e45e2ec67b77cull, 0x3fefffc000155552ull, 0x3c8f4a01a0196daeull,
    0x3f8fffaaaaeeeed5ull, 0xbc02ab639a9f0776ull, 0x3fefff000155549full, 0x3c828a28a03a5ef3ull,
    0x3f97ff7001033255ull, 0x3bfefe2b51527336ull, 0x3feffdc006bff7e6ull, 0x3c8ae6dae86977bdull,
    0x3f9ffeaaaeeee86full, 0xbc3cd406fb224ae2ull, 0x3feffc00155527d3ull, 0xbc83b54492d89b5bull,
    0x3fa3feb2b12d45d5ull, 0x3c34ec54203d1c11ull, 0x3feff9c03414a7baull, 0x3c6991f4be6c59bfull,
    0x3fa7fdc01032fba9ull, 0xbc4599bdf46e997aull, 0x3feff7006bfdf99full, 0xbc78b3b560648d5full,
    0x3fabfc6d78586dacull, 0x3c18e4fd03dbf236ull, 0x3feff3c0c8103a31ull, 0x3c74856dbddc0e66ull,
    0x3faffaaaeeed4edbull, 0xbc42d16d32684b69ull, 0x3feff0015549f4d3ull, 0x3c8328387b99426full,
    0x3fb1fc343d808befull, 0xbc5f3d32e6f3be4full, 0x3fefebc222a8ef9full, 0x3c57934934f54c77ull,
    0x3fb3facb12d1755bull, 0xbc5921915299468bull, 0x3fefe7034129ef6full, 0xbc6cbf4337c96f97ull,
    0x3fb5f911fd10b737ull, 0xbc50184f02be9102ull, 0x3fefe1c4c3c873ebull, 0xbc35a9c9057c4a02ull,
    0x3fb7f701032550e4ull, 0x3c3afc2d1800501aull, 0x3fefdc06bf7e6b9bull, 0x3c831902b535f8dbull,
    0x3fb9f4902d55d1f9ull, 0x3c52696d7eac1dc1ull, 0x3fefd5c94b43e000ull, 0xbc62e768cb4f92f9ull,
    0x3fbbf1b78568391dull, 0x3c5e91841dea4cc8ull, 0x3fefcf0c800e99b1ull, 0x3c6ea3d786d186acull,
    0x3fbdee6f16c1cce6ull, 0xbc450f8e2fb71673ull, 0x3fefc7d078d1bc88ull, 0x3c8075d2447db685ull,
    0x3fbfeaaeee86ee36ull, 0xbc4afcb2bcc6f03bull, 0x3fefc015527d5bd3ull, 0x3c8b68f35094efb8ull,
    0x3fc0f3378ddd71d1ull, 0x3c6d8468724f0f9eull, 0x3fefb7db2bfe0695ull, 0x3c821dadf4f65ab1ull,
    0x3fc1f0d3d7afceafull, 0xbc66ef95099769a5ull, 0x3fefaf22263c4bd3ull, 0xbc552ace133a2769ull,
    0x3fc2ee285e4ab88full, 0xbc6e4d0f05dee058ull, 0x3fefa5ea641c36f2ull, 0x3c404da6ed17cc7cull,
    0x3fc3eb312c5d66cbull, 0x3c647d666b66cb91ull, 0x3fef9c340a7cc428ull, 0x3c8c5b6b063b7462ull,
    0x3fc4e7ea4dc5f27bull, 0x3c5949db2ac072fcull, 0x3fef91ff40374d01ull, 0xbc67d03f4d3a9e4cull,
    0x3fc5e44fcfa126f3ull, 0xbc66f443063f89b6ull, 0x3fef874c2e1eecf6ull, 0xbc8c6514e1332b16ull,
    0x3fc6e05dc05a4d4cull, 0xbbd32c5c8b81c919ull, 0x3fef7c1afeffde24ull, 0xbc78f55bc47540b1ull,
    0x3fc7dc102fbaf2b5ull, 0x3c45ab50e23c97c3ull, 0x3fef706bdf9ece1cull, 0xbc8698c80c36dcb4ull,
    0x3fc8d7632efaa944ull, 0xbc620fa262cbb953ull, 0x3fef643efeb82acdull, 0x3c76b00ac1fe28acull,
    0x3fc9d252d0cec312ull, 0x3c59c43d80b1137dull, 0x3fef57948cff6797ull, 0x3c6e3a0d3e03b1d4ull,
    0x3fcaccdb297a0765ull, 0xbc59883b57d6cdeaull, 0x3fef4a6cbd1e3a79ull, 0x3c813df0edaebb57ull,
    0x3fcbc6f84edc6199ull, 0x3c69c1a56a7b0cabull, 0x3fef3cc7c3b3d16eull, 0xbc621a3ad28a3494ull,
    0x3fccc0a6588289a3ull, 0xbc6868d09bc87c6bull, 0x3fef2ea5d753ffedull, 0x3c8cc4215f56d583ull,
    0x3fcdb9e15fb5a5d0ull, 0xbc632e20d6cc6fc2ull, 0x3fef20073086649full, 0x3c7b940416c1984bull,
    0x3fceb2a57f8ae5a3ull, 0xbc60be06af572cebull, 0x3fef10ec09c5873bull, 0x3c8d9072762c1283ull,
    0x3fcfaaeed4f31577ull, 0xbc615d88508e32b8ull, 0x3fef01549f7deea1ull, 0x3c8d3c1e99e5cafdull,
    0x3fd0515cbf65155cull, 0xbc79b8c29dfd8ec7ull, 0x3feef141300d2f26ull, 0xbc82aa1b08ded372ull,
    0x3fd0cd00cef36436ull, 0xbc79fb0a0c93e2b4ull, 0x3feee0b1fbc0f11cull, 0xbc4bfd2380bbc3b1ull,
    0x3fd14861aa94ddebull, 0xbc6be881b5b615a4ull, 0x3feecfa744d5efa1ull, 0xbc556d0a4af541d0ull,
    0x3fd1c37d64c6b876ull, 0x3c746076fe0dcff4ull, 0x3feebe214f76efa8ull, 0xbc802f9f12ba543eull,
    0x3fd23e52111aaf36ull, 0xbc74f080334eff18ull, 0x3feeac2061bbaf4full, 0x3c62c1d53e94658dull,
    0x3fd2b8ddc43eb49full, 0x3c61553899f2d807ull, 0x3fee99a4c3a7cd83ull, 0xbc82264b1bc53ce8ull,
    0x3fd3331e94049f87ull, 0x3c7e0cb6b40c302cull, 0x3fee86aebf29a9edull, 0x3c89397afdbb58a7ull,
    0x3fd3ad129769d3d8ull, 0x3c003d550487839aull, 0x3fee733ea0193d40ull, 0xbc86428b3546ce13ull,
    0x3fd426b7e69ee697ull, 0xbc7f09c75705c59full, 0x3fee5f54b436e9d0ull, 0x3c87eb0fd02fc8bcull,
    0x3fd4a00c9b0f3d20ull, 0x3c7823ba6bb08eadull, 0x3fee4af14b2a449cull, 0xbc868ca02e8a6833ull,
    0x3fd5190ecf68a77aull, 0x3c7b357155eef0f3ull, 0x3fee3614b680d6a5ull, 0xbc727793aa015237ull,
    0x3fd591bc9fa2f597ull, 0x3c67c74bac3fe0cbull, 0x3fee20bf49acd6c1ull, 0xbc5660aec7ef636bull,
    0x3fd60a1429078775ull, 0x3c5b1fd80ba89133ull, 0x3fee0af15a03dbceull, 0x3c5fe8e702771ae6ull,
    0x3fd682138a38d7f7ull, 0xbc7d889202444aadull, 0x3fedf4ab3ebd875eull, 0xbc8e2d8a7e6736c4ull,
    0x3fd6f9b8e33a0255ull, 0x3c742bc14ee9da0dull, 0x3feddded50f228d6ull, 0xbc6e80c8d42ba2bfull,
    0x3fd7710255764214ull, 0xbc66ead7314bb6ceull, 0x3fedc6b7eb995912ull, 0x3c54b364776dcd35ull,
    0x3fd7e7ee03c86d4eull, 0xbc7b63bcdabf5af2ull, 0x3fedaf0b6b888e83ull, 0x3c8a249e2b5e5ceaull,
    0x3fd85e7a12826949ull, 0x3c78a40e9b5face0ull, 0x3fed96e82f71a9dcull, 0x3c8ff61bd5d2039dull,
    0x3fd8d4a4a774992full, 0x3c744a02ea766326ull, 0x3fed7e4e97e17b4aull, 0xbc63b770352bed94ull,
    0x3fd94a6be9f546c5ull, 0xbc769ce13e683f58ull, 0x3fed653f073e4040ull, 0xbc876236434bec37ull,
    0x3fd9bfce02e80510ull, 0x3c709e39a320b0a4ull, 0x3fed4bb9e1c619e0ull, 0x3c8f34bb77858f61ull,
    0x3fda34c91cc50ccaull, 0xbc5a310e3b50cecdull, 0x3fed31bf8d8d7c06ull, 0x3c7e60dd3089cbddull,
    0x3fdaa95b63a09277ull, 0xbc66293eb13c0381ull, 0x3fed1750727d94f0ull, 0x3c80d52b1ec1a48eull,
    0x3fdb1d8305321617ull, 0xbc7ae242cb99f519ull, 0x3fecfc6cfa52ad9full, 0x3c88b5b5508f2a0dull,
    0x3fdb913e30dbac43ull, 0xbc7e38ad2f6c3ff1ull, 0x3fece115909a82e5ull, 0x3c81f139bb31109aull,
    0x3fdc048b17b140a3ull, 0x3c619fe6757e9fa7ull, 0x3fecc54aa2b2972eull, 0x3c64ee162ba83a98ull,
    0x3fdc7767ec7fd19eull, 0xbc5eb14d1a3d5826ull, 0x3feca90c9fc67d0bull, 0xbc646a81485e3462ull,
    0x3fdce9d2e3d4a51full, 0xbc62fc8a12dae298ull, 0x3fec8c5bf8ce1a84ull, 0x3c7ab3d1a1590123ull,
    0x3fdd5bca34047661ull, 0x3c728a44a75fc29cull, 0x3fec6f39208be53bull, 0xbc8741dbfbaadb42ull,
    0x3fddcd4c15329c9aull, 0x3c70d4c6e171fd9aull, 0x3fec51a48b8b175eull, 0xbc61bbb43b9aa880ull,
    0x3fde3e56c1582a69ull, 0xbc50a4821099f88full, 0x3fec339eb01ddd81ull, 0xbc8caaf5ee82c5c0ull,
    0x3fdeaee8744b05f0ull, 0xbc5789b43c9b027dull, 0x3fec1528065b7d50ull, 0xbc8892111312e828ull,
    0x3fdf1eff6bc4f97bull, 0x3c717212f8a7525cull, 0x3febf641081e7536ull, 0x3c8b7bd71628a9a1ull,
    0x3fdf8e99e76abc97ull, 0x3c59d950af2d00a3ull, 0x3febd6ea310294f5ull, 0x3c731bbcc88c109dull,
    0x3fdffdb628d2f57aull, 0x3c6f4a992e905b6aull, 0x3febb723fe630f32ull, 0x3c772bd2452d0a39ull,
    0x3fe0362939c69955ull, 0xbc82d8cd78397b01ull, 0x3feb96eeef58840eull, 0x3c545a3cc78fade0ull,
    0x3fe06d3686946e5bull, 0x3c83f5ae4538ff1bull, 0x3feb764b84b704c2ull, 0xbc8f5848c21b389bull,
    0x3fe0a4021e9e1001ull, 0xbc86f643a13914f6ull, 0x3feb553a410c104eull, 0x3c58ff7947027a15ull,
    0x3fe0da8b26b5672eull, 0xbc8a58def0bee909ull, 0x3feb33bba89c8948ull, 0x3c8ea6a51d1f6ca9ull,
    0x3fe110d0c4b69c3bull, 0x3c8d918998809981ull, 0x3feb11d04162a4c6ull, 0x3c71dd561efbc0c2ull,
    0x3fe146d21f8b7f82ull, 0x3c7bf9535e2739a8ull, 0x3feaef78930bd275ull, 0xbc7f836279746f94ull,
    0x3fe17c8e5f2eedb0ull, 0x3c635e57102e2488ull, 0x3feaccb526f69de5ull, 0x3c88fb6a8dd6b6ccull,
    0x3fe1b204acb02fddull, 0xbc5f190c70cbb5feull, 0x3feaa98688308913ull, 0xbc0b83d607cd5072ull,
    0x3fe1e7343236574cull, 0x3c722a3fa4f41d5aull, 0x3fea85ed4373e02dull, 0x3c69be06385ec792ull,
    0x3fe21c1c1b0394cfull, 0x3c5e5b324b23aa31ull, 0x3fea61e9e72586afull, 0x3c858330e2fd453full,
    0x3fe250bb93788bbbull, 0x3c7ea3d02457bcceull, 0x3fea3d7d0352bdcfull, 0xbc868dbaeca19669ull,
    0x3fe28511c917a067ull, 0xbc801df1d9a16b70ull, 0x3fea18a729aee445ull, 0x3c395e25736c0357ull,
    0x3fe2b91dea88421eull, 0xbc8fa371db216ab0ull, 0x3fe9f368ed912f85ull, 0xbc81d200c5791606ull,
    0x3fe2ecdf279a3082ull, 0x3c8d3557e0e7e37eull, 0x3fe9cdc2e3f25e5cull, 0x3c83f99112993f62ull,
    0x3fe32054b148bc4full, 0x3c8f6b42095a135bull, 0x3fe9a7b5a36a6514ull, 0x3c8722cfcc9fa7a9ull,
    0x3fe3537db9be0367ull, 0x3c6b327e7af040f0ull, 0x3fe98141c42e1310ull, 0x3c8d1ff80488f08dull,
    0x3fe386597456282bull, 0xbc710fada93b07a8ull, 0x3fe95a67e00cb1fdull, 0xbc80befda21f862dull,
    0x3fe3b8e715a2840aull, 0xbc797653a7d2f07aull, 0x3fe93328926d9e92ull, 0xbc8bb77003600cdaull,
    0x3fe3eb25d36cd53aull, 0xbc5be570e1570fc0ull, 0x3fe90b84784ddaf7ull, 0xbc70feb10ab93b87ull,
    0x3fe41d14e4ba6790ull, 0x3c84608fd287ecf5ull, 0x3fe8e37c303d9ad1ull, 0xbc6463a4b53d4bf8ull,
    0x3fe44eb381cf386bull, 0xbc83ed6c1e6a5505ull, 0x3fe8bb105a5dc900ull, 0x3c8863e03e9474c1ull,
    0x3fe48000e431159full, 0xbc8b194a7463ed10ull, 0x3fe89241985d871full, 0x3c8c48d9c413ed84ull,
    0x3fe4b0fc46aab761ull, 0x3c20da05738cc59cull, 0x3fe869108d77a6c6ull, 0x3c7338ffe2bfe9ddull,
    0x3fe4e1a4e54ed51bull, 0xbc8a492f89b7c76aull, 0x3fe83f7dde701ca0ull, 0xbc4152cf609bc6e8ull,
    0x3fe511f9fd7b351cull, 0xbc85c0e861c48831ull, 0x3fe8158a31916d5dull, 0xbc6de8b90b8228deull,
    0x3fe541facddbb724ull, 0x3c7232c28520d391ull, 0x3fe7eb362eaa1488ull, 0x3c5a1d65a4a5959full,
    0x3fe571a6966d59b3ull, 0x3c5c843b4d0fb197ull, 0x3fe7c0827f09e54full, 0xbc6c73d6d72aee68ull,
    0x3fe5a0fc98813a12ull, 0xbc8d82e2b7d4227bull, 0x3fe7956fcd7f6543ull, 0xbc8ab276e9d45ae4ull,
    0x3fe5cffc16bf8f0dull, 0x3c896cb370eb578aull, 0x3fe769fec655211full, 0xbc6827d5cf8c68c5ull,
    0x3fe5fea4552a9e57ull, 0x3c80b6cef7ee20b7ull, 0x3fe73e30174efba1ull, 0xbc65d3ae3d94ad5full,
    0x3fe62cf49921ac79ull, 0xbc8edd9855b6241aull, 0x3fe712046fa77678ull, 0x3c8425b0a5029c81ull,
    0x3fe65aec2963e755ull, 0x3c8126f96b71053cull, 0x3fe6e57c800cf55eull, 0x3c860286dedbd0a6ull,
    0x3fe6888a4e134b2full, 0xbc86b7d37644d5e6ull, 0x3fe6b898fa9efb5dull, 0x3c715ac786ccf4b2ull,
    0x3fe6b5ce50b7821aull, 0xbc65d5158f702e0full, 0x3fe68b5a92eb6253ull, 0xbc89a91ad985f89cull,
    0x3fe6e2b77c40bde1ull, 0xbc70e729857fad53ull, 0x3fe65dc1fdeb8cbaull, 0xbc597c1b47337c77ull,
    0x3fe70f451d0a8c40ull, 0x3c697ede3885770dull, 0x3fe62fcff20191c7ull, 0x3c6d9143895756efull,
    0x3fe73b7680dea578ull, 0xbc72248306dc12a2ull, 0x3fe6018526f563dfull, 0x3c846ca5e0e432d0ull,
    0x3fe7674af6f7b524ull, 0x3c7e9d3f94ac84a8ull, 0x3fe5d2e255f1f17aull, 0x3c80314104c8892bull,
    0x3fe792c1d0041d52ull, 0xbc8abf05eeb354ebull, 0x3fe5a3e839824077ull, 0x3c8428aa2759be62ull,
    0x3fe7bdda5e28b3c2ull, 0x3c4ad1197ccd0392ull, 0x3fe574978d8e83f2ull, 0x3c8f4714af282d23ull,
    0x3fe7e893f5037959ull, 0x3c80eefbaa650c4cull, 0x3fe544f10f592ca5ull, 0xbc8e7ae8e6c7a62full,
    0x3fe812ede9ae4ba4ull, 0xbc87830adf402ddaull, 0x3fe514f57d7bf3daull, 0x3c747a108073c259ull,
};

namespace libm_detail {

constexpr double kSn3 = -1.66666666666664880952546298448555E-01, kSn5 = 8.33333214285722277379541354343671E-03,
                 kCs2 = 4.99999999999999999999950396842453E-01, kCs4 = -4.16666666666664434524222570944589E-02,
                 kCs6 = 1.38888874007937613028114285595617E-03;
constexpr double kS1 = -0x1.5555555555555p-3, kS2 = 0x1.1111111110ECEp-7, kS3 = -0.00019841269834414642,
                 kS4 = 2.755729806860771e-06, kS5 = -2.5022014848318398e-08;
constexpr double kBig = 0x1.8p45;                                       // ulp 2^-7: rounds |x| to i/128
constexpr double kHp0 = 0x1.921FB54442D18p0, kHp1 = 0x1.1A62633145C07p-54;   // pi/2 = kHp0 + kHp1
constexpr double kHpInv = 0x1.45F306DC9C883p-1, kToInt = 0x1.8p52;
constexpr double kMp1 = 0x1.921FB58p0, kMp2 = -0x1.DDE973Cp-27, kPp3 = -0x1.CB3B398p-55,
                 kPp4 = -0x1.d747f23e32ed7p-83;                          // pi/2 in four pieces

__device__ inline double tab(int k) { return __longlong_as_double((long long)kSinCosTable[k]); }

// sin(a + da) for |a| < 0.126: a - a^3/6 + ... with the tail's first-order term.
__device__ inline double taylor_sin(double xx, double a, double da)
{
#pragma clang fp contract(off)
    const double poly = ((((kS5 * xx + kS4) * xx + kS3) * xx + kS2) * xx) + kS1;
    const double t = ((poly * a - 0.5 * da) * xx) + da;
    return a + t;
}

// cos(x + dx) for |x| < 0.855469 (x + dx a double-double).
__device__ inline double do_cos(double x, double dx)
{
#pragma clang fp contract(off)
    if (x < 0) dx = -dx;
    const double u = kBig + fabs(x);
    x = fabs(x) - (u - kBig) + dx;
    const double xx = x * x;
    const double s = x + x * xx * (kSn3 + xx * kSn5);
    const double c = xx * (kCs2 + xx * (kCs4 + xx * kCs6));
    const int k = (int)((uint32_t)__double_as_longlong(u)) << 2;
    const double sn = tab(k), ssn = tab(k + 1), cs = tab(k + 2), ccs = tab(k + 3);
    const double cor = (ccs - s * ssn - cs * c) - sn * s;
    return cs + cor;
}

// sin(x + dx) for |x| < 0.855469.
__device__ inline double do_sin(double x, double dx)
{
#pragma clang fp contract(off)
    const double xold = x;
    if (fabs(x) < 0.126) return taylor_sin(x * x, x, dx);
    if (x <= 0) dx = -dx;
    const double u = kBig + fabs(x);
    x = fabs(x) - (u - kBig);
    const double xx = x * x;
    const double s = x + (dx + x * xx * (kSn3 + xx * kSn5));
    const double c = x * dx + xx * (kCs2 + xx * (kCs4 + xx * kCs6));
    const int k = (int)((uint32_t)__double_as_longlong(u)) << 2;
    const double sn = tab(k), ssn = tab(k + 1), cs = tab(k + 2), ccs = tab(k + 3);
    const double cor = (ssn + s * ccs - sn * c) + cs * s;
    return copysign(sn + cor, xold);
}

// x = n pi/2 + a + da for |x| < 105414350; returns n mod 4.
__device__ inline int reduce(double x, double& a, double& da)
{
#pragma clang fp contract(off)
    const double t = x * kHpInv + kToInt;
    const double xn = t - kToInt;
    const int n = (int)((uint64_t)__double_as_longlong(t) & 3u);
    const double y = (x - xn * kMp1) - xn * kMp2;
    double t1 = xn * kPp3;
    const double t2 = y - t1;
    double db = (y - t2) - t1;
    t1 = xn * kPp4;
    const double b = t2 - t1;
    db += (t2 - b) - t1;
    a = b;
    da = db;
    return n;
}

__device__ inline double do_sincos(double a, double da, int n)
{
    const double r = (n & 1) ? do_cos(a, da) : do_sin(a, da);
    return (n & 2) ? -r : r;
}

} // namespace libm_detail

__device__ inline void libm_sincos(double x, double& sinx, double& cosx)
{
#pragma clang fp contract(off)
    using namespace libm_detail;
    const uint32_t k = (uint32_t)((uint64_t)__double_as_longlong(x) >> 32) & 0x7fffffffu;
    if (k < 0x3e400000u) {            // |x| < 2^-27
        sinx = x;
        cosx = 1.0;
    } else if (k < 0x3feb6000u) {     // |x| < 0.855469
        sinx = do_sin(x, 0.0);
        cosx = do_cos(x, 0.0);
    } else if (k < 0x400368fdu) {     // |x| < 2.426265
        const double y = kHp0 - fabs(x);
        const double a = y + kHp1;
        const double da = (y - a) + kHp1;
        sinx = copysign(do_cos(y, kHp1), x);
        cosx = do_sin(a, da);
    } else if (k < 0x419921fbu) {     // |x| < 105414350
        double a, da;
        const int n = reduce(x, a, da);
        sinx = do_sincos(a, da, n);
        cosx = do_sincos(a, da, n + 1);
    } else if (k >= 0x7ff00000u) {    // inf, nan
        sinx = cosx = x / x;
    } else {
        sinx = sin(x);
        cosx = cos(x);
    }
}

} // namespace pt
