// Table of the decoder's exp (the table-driven exp of Arm Optimized Routines,
// N = 128): per j, the tail T[j] of 2^(j/128) = H[j] (1 + T[j]) and the bits of
// H[j] minus j << 45, generated with 80-digit decimal arithmetic by
// make_exp_table.py; the reduction constants and polynomial coefficients of
// that design.
#pragma once

#include <stdint.h>

namespace sg {

constexpr int kExpTableBits = 7;
constexpr double kExpInvLn2N = 0x1.71547652b82fep0 * 128; // 128 / ln 2
constexpr double kExpNegLn2HiN = -0x1.62e42fefa0000p-8;  // -ln 2 / 128, high part
constexpr double kExpNegLn2LoN = -0x1.cf79abc9e3b3ap-47; // -ln 2 / 128, low part
constexpr double kExpShift = 0x1.8p52;                   // round-to-integer shifter
// exp(r) - 1 - T ~ r + r^2 (C2 + r C3) + r^4 (C4 + r C5), |r| <= ln 2 / 256
constexpr double kExpC2 = 0x1.ffffffffffdbdp-2;
constexpr double kExpC3 = 0x1.555555555543cp-3;
constexpr double kExpC4 = 0x1.55555cf172b91p-5;
constexpr double kExpC5 = 0x1.1111167a4d017p-7;

struct ExpTail {
    double tail;
    uint64_t sbits;
};

constexpr ExpTail kExp2Table[128] = {
    {0x0.0p+0, 0x3ff0000000000000ull}, // 2^(0/128)
    {0x1.b3b4f1a88bf6ep-54, 0x3feff63da9fb3335ull}, // 2^(1/128)
    {-0x1.160139cd8dc5dp-56, 0x3fefec9a3e778061ull}, // 2^(2/128)
    {-0x1.05e7a108766d1p-54, 0x3fefe315e86e7f85ull}, // 2^(3/128)
    {0x1.cd2523567f613p-55, 0x3fefd9b0d3158574ull}, // 2^(4/128)
    {-0x1.bce8023f98efap-55, 0x3fefd06b29ddf6deull}, // 2^(5/128)
    {0x1.0f74e61e6c861p-57, 0x3fefc74518759bc8ull}, // 2^(6/128)
    {0x1.0a3e45b33d399p-54, 0x3fefbe3ecac6f383ull}, // 2^(7/128)
    {0x1.79aa65d837b6dp-54, 0x3fefb5586cf9890full}, // 2^(8/128)
    {0x1.eb51a92fdeffcp-55, 0x3fefac922b7247f7ull}, // 2^(9/128)
    {0x1.ebe3d702f9cd1p-60, 0x3fefa3ec32d3d1a2ull}, // 2^(10/128)
    {-0x1.a033489906e0bp-57, 0x3fef9b66affed31bull}, // 2^(11/128)
    {-0x1.556522a2fbd0ep-54, 0x3fef9301d0125b51ull}, // 2^(12/128)
    {-0x1.080ef8c4eea55p-58, 0x3fef8abdc06c31ccull}, // 2^(13/128)
    {-0x1.1c923b9d5f416p-54, 0x3fef829aaea92de0ull}, // 2^(14/128)
    {0x1.0d3e3e95c55afp-55, 0x3fef7a98c8a58e51ull}, // 2^(15/128)
    {-0x1.01b15eaa59348p-55, 0x3fef72b83c7d517bull}, // 2^(16/128)
    {-0x1.f1ff055de323dp-55, 0x3fef6af9388c8deaull}, // 2^(17/128)
    {0x1.b898c3f1353bfp-55, 0x3fef635beb6fcb75ull}, // 2^(18/128)
    {-0x1.6d99c7611eb26p-54, 0x3fef5be084045cd4ull}, // 2^(19/128)
    {0x1.aecf73e3a2f60p-54, 0x3fef54873168b9aaull}, // 2^(20/128)
    {-0x1.fe782cb86389dp-55, 0x3fef4d5022fcd91dull}, // 2^(21/128)
    {0x1.a6f4144a6c38dp-55, 0x3fef463b88628cd6ull}, // 2^(22/128)
    {0x1.07a05b0e4047dp-55, 0x3fef3f49917ddc96ull}, // 2^(23/128)
    {0x1.68efde3a8a894p-54, 0x3fef387a6e756238ull}, // 2^(24/128)
    {0x1.75e18f274487dp-55, 0x3fef31ce4fb2a63full}, // 2^(25/128)
    {0x1.0472b981fe7f2p-55, 0x3fef2b4565e27cddull}, // 2^(26/128)
    {-0x1.6b87b3f71085ep-54, 0x3fef24dfe1f56381ull}, // 2^(27/128)
    {0x1.2f7e16d09ab31p-55, 0x3fef1e9df51fdee1ull}, // 2^(28/128)
    {-0x1.d219b1a6fbffap-60, 0x3fef187fd0dad990ull}, // 2^(29/128)
    {0x1.b3782720c0ab4p-55, 0x3fef1285a6e4030bull}, // 2^(30/128)
    {0x1.e149289cecb8fp-57, 0x3fef0cafa93e2f56ull}, // 2^(31/128)
    {0x1.34d754db0abb6p-55, 0x3fef06fe0a31b715ull}, // 2^(32/128)
    {0x1.64201e2ac744cp-55, 0x3fef0170fc4cd831ull}, // 2^(33/128)
    {0x1.fdd395dd3f84ap-55, 0x3feefc08b26416ffull}, // 2^(34/128)
    {-0x1.6a3803b8e5b04p-55, 0x3feef6c55f929ff1ull}, // 2^(35/128)
    {-0x1.24aedcc4b5068p-54, 0x3feef1a7373aa9cbull}, // 2^(36/128)
    {-0x1.907f81b512d8ep-54, 0x3feeecae6d05d866ull}, // 2^(37/128)
    {-0x1.1d1e83e9436d2p-56, 0x3feee7db34e59ff7ull}, // 2^(38/128)
    {-0x1.91919b3ce1b15p-54, 0x3feee32dc313a8e5ull}, // 2^(39/128)
    {0x1.59f48a72a4c6dp-55, 0x3feedea64c123422ull}, // 2^(40/128)
    {-0x1.312607a28698ap-54, 0x3feeda4504ac801cull}, // 2^(41/128)
    {-0x1.8a78f4817895bp-58, 0x3feed60a21f72e2aull}, // 2^(42/128)
    {-0x1.c2c9b67499a1bp-56, 0x3feed1f5d950a897ull}, // 2^(43/128)
    {0x1.363ed60c2ac11p-59, 0x3feece086061892dull}, // 2^(44/128)
    {0x1.666093b0664efp-54, 0x3feeca41ed1d0057ull}, // 2^(45/128)
    {0x1.ecce1daa10379p-57, 0x3feec6a2b5c13cd0ull}, // 2^(46/128)
    {0x1.3ff8e3f0f1230p-54, 0x3feec32af0d7d3deull}, // 2^(47/128)
    {0x1.690cebb7aafb0p-56, 0x3feebfdad5362a27ull}, // 2^(48/128)
    {0x1.31dbdeb54e077p-54, 0x3feebcb299fddd0dull}, // 2^(49/128)
    {-0x1.f94340071a38ep-55, 0x3feeb9b2769d2ca7ull}, // 2^(50/128)
    {-0x1.7deccdc93a349p-55, 0x3feeb6daa2cf6642ull}, // 2^(51/128)
    {-0x1.8dec6bd0f385fp-56, 0x3feeb42b569d4f82ull}, // 2^(52/128)
    {-0x1.61246ec7b5cf6p-55, 0x3feeb1a4ca5d920full}, // 2^(53/128)
    {0x1.3350518fdd78ep-54, 0x3feeaf4736b527daull}, // 2^(54/128)
    {0x1.b98b72f8a9b05p-56, 0x3feead12d497c7fdull}, // 2^(55/128)
    {0x1.063e1e21c5409p-54, 0x3feeab07dd485429ull}, // 2^(56/128)
    {0x1.4c7855019c6eap-60, 0x3feea9268a5946b7ull}, // 2^(57/128)
    {0x1.432e62b64c035p-54, 0x3feea76f15ad2148ull}, // 2^(58/128)
    {-0x1.ce44a6199769fp-55, 0x3feea5e1b976dc09ull}, // 2^(59/128)
    {-0x1.c33c53bef4da8p-55, 0x3feea47eb03a5585ull}, // 2^(60/128)
    {-0x1.45378892be9aep-55, 0x3feea34634ccc320ull}, // 2^(61/128)
    {-0x1.3cedd78565858p-54, 0x3feea23882552225ull}, // 2^(62/128)
    {0x1.710aa807e1964p-58, 0x3feea155d44ca973ull}, // 2^(63/128)
    {-0x1.3b3efbf5e2228p-54, 0x3feea09e667f3bcdull}, // 2^(64/128)
    {-0x1.a12ad8734b982p-57, 0x3feea012750bdabfull}, // 2^(65/128)
    {-0x1.367efb86da9eep-57, 0x3fee9fb23c651a2full}, // 2^(66/128)
    {-0x1.0dc3d54e08851p-55, 0x3fee9f7df9519484ull}, // 2^(67/128)
    {-0x1.81f647e5a3ecfp-56, 0x3fee9f75e8ec5f74ull}, // 2^(68/128)
    {-0x1.6ee4ac08b7db0p-55, 0x3fee9f9a48a58174ull}, // 2^(69/128)
    {-0x1.619321e55e68ap-55, 0x3fee9feb564267c9ull}, // 2^(70/128)
    {0x1.09ccb5e09d4d3p-54, 0x3feea0694fde5d3full}, // 2^(71/128)
    {-0x1.b32dcb94da51dp-56, 0x3feea11473eb0187ull}, // 2^(72/128)
    {0x1.4ecfd5467c06bp-54, 0x3feea1ed0130c132ull}, // 2^(73/128)
    {0x1.5ebe1abd66c55p-57, 0x3feea2f336cf4e62ull}, // 2^(74/128)
    {-0x1.8a1c52fb3cf42p-55, 0x3feea427543e1a12ull}, // 2^(75/128)
    {-0x1.369b6f13b3734p-54, 0x3feea589994cce13ull}, // 2^(76/128)
    {-0x1.05e843a19ff1ep-55, 0x3feea71a4623c7adull}, // 2^(77/128)
    {-0x1.4d450d872576ep-54, 0x3feea8d99b4492edull}, // 2^(78/128)
    {0x1.0ad675b0e8a00p-54, 0x3feeaac7d98a6699ull}, // 2^(79/128)
    {0x1.db72fc1f0eab4p-55, 0x3feeace5422aa0dbull}, // 2^(80/128)
    {-0x1.5b6609cc5e7ffp-57, 0x3feeaf3216b5448cull}, // 2^(81/128)
    {0x1.bf68359f35f44p-56, 0x3feeb1ae99157736ull}, // 2^(82/128)
    {-0x1.3091fa71e3d83p-54, 0x3feeb45b0b91ffc6ull}, // 2^(83/128)
    {-0x1.da9b88b6c1e29p-58, 0x3feeb737b0cdc5e5ull}, // 2^(84/128)
    {-0x1.c23f97c90b959p-57, 0x3feeba44cbc8520full}, // 2^(85/128)
    {-0x1.2434322f4f9aap-54, 0x3feebd829fde4e50ull}, // 2^(86/128)
    {-0x1.5ca6cd7668e4bp-55, 0x3feec0f170ca07baull}, // 2^(87/128)
    {0x1.1affc2b91ce27p-56, 0x3feec49182a3f090ull}, // 2^(88/128)
    {0x1.dd235e10a73bbp-57, 0x3feec86319e32323ull}, // 2^(89/128)
    {-0x1.7c50422622263p-55, 0x3feecc667b5de565ull}, // 2^(90/128)
    {0x1.b1c86e3e231d5p-55, 0x3feed09bec4a2d33ull}, // 2^(91/128)
    {-0x1.1bbd1d3bcbb15p-54, 0x3feed503b23e255dull}, // 2^(92/128)
    {0x1.0cc319cee31d2p-54, 0x3feed99e1330b358ull}, // 2^(93/128)
    {0x1.469846e735ab3p-55, 0x3feede6b5579fdbfull}, // 2^(94/128)
    {-0x1.2dfcd978e9db4p-55, 0x3feee36bbfd3f37aull}, // 2^(95/128)
    {0x1.c1a7792cb3387p-55, 0x3feee89f995ad3adull}, // 2^(96/128)
    {-0x1.07b8f4ad1d9fap-54, 0x3feeee07298db666ull}, // 2^(97/128)
    {-0x1.5c3d956dcaebap-58, 0x3feef3a2b84f15fbull}, // 2^(98/128)
    {-0x1.0a40e3da6f640p-54, 0x3feef9728de5593aull}, // 2^(99/128)
    {-0x1.8d6f438ad9334p-57, 0x3feeff76f2fb5e47ull}, // 2^(100/128)
    {-0x1.1eee26b588a35p-54, 0x3fef05b030a1064aull}, // 2^(101/128)
    {0x1.4ffd70a5fddcdp-56, 0x3fef0c1e904bc1d2ull}, // 2^(102/128)
    {-0x1.1bdfbfa9298acp-54, 0x3fef12c25bd71e09ull}, // 2^(103/128)
    {0x1.36eae30af0cb3p-56, 0x3fef199bdd85529cull}, // 2^(104/128)
    {0x1.ee3325c9ffd94p-55, 0x3fef20ab5fffd07aull}, // 2^(105/128)
    {0x1.4e08fd10959acp-55, 0x3fef27f12e57d14bull}, // 2^(106/128)
    {0x1.3cdaf384e1a67p-57, 0x3fef2f6d9406e7b5ull}, // 2^(107/128)
    {0x1.76b2c6c921968p-57, 0x3fef3720dcef9069ull}, // 2^(108/128)
    {-0x1.08a1883ccb5d2p-55, 0x3fef3f0b555dc3faull}, // 2^(109/128)
    {-0x1.fad5d3ffffa6fp-55, 0x3fef472d4a07897cull}, // 2^(110/128)
    {-0x1.00dae3875a949p-54, 0x3fef4f87080d89f2ull}, // 2^(111/128)
    {0x1.4a385a63d07a7p-56, 0x3fef5818dcfba487ull}, // 2^(112/128)
    {-0x1.2919e2040220fp-55, 0x3fef60e316c98398ull}, // 2^(113/128)
    {0x1.e5a50d5c192acp-55, 0x3fef69e603db3285ull}, // 2^(114/128)
    {0x1.43a59ac016b4bp-55, 0x3fef7321f301b460ull}, // 2^(115/128)
    {-0x1.2d52107b43e1fp-55, 0x3fef7c97337b9b5full}, // 2^(116/128)
    {-0x1.92ab93b470dc9p-55, 0x3fef864614f5a129ull}, // 2^(117/128)
    {0x1.4b604603a88d3p-56, 0x3fef902ee78b3ff6ull}, // 2^(118/128)
    {0x1.3c5ec519d7271p-55, 0x3fef9a51fbc74c83ull}, // 2^(119/128)
    {-0x1.ff7128fd391f0p-55, 0x3fefa4afa2a490daull}, // 2^(120/128)
    {-0x1.dae98e223747dp-55, 0x3fefaf482d8e67f1ull}, // 2^(121/128)
    {0x1.ec3bc41aa2008p-55, 0x3fefba1bee615a27ull}, // 2^(122/128)
    {0x1.42b94c3a9eb32p-55, 0x3fefc52b376bba97ull}, // 2^(123/128)
    {0x1.a64a931d185eep-55, 0x3fefd0765b6e4540ull}, // 2^(124/128)
    {-0x1.e37bae43be3edp-55, 0x3fefdbfdad9cbe14ull}, // 2^(125/128)
    {0x1.7893b4d91cd9dp-56, 0x3fefe7c1819e90d8ull}, // 2^(126/128)
    {0x1.305c14160cc89p-58, 0x3feff3c22b8f71f1ull}, // 2^(127/128)
};

} // namespace sg
