// Generated by mat_amd/autotune_batched.py on one MI355X: panel width of
// the batched uneven QR per m bin, the faster of 8 and 16 on 1500 matrices
// per bin (n uniform in [8, m]), median of 3 hipGraph replays.
//
//   float  m in [ 16,  64]: nb 8 0.4561 ms, nb 16 0.4823 ms
//   float  m in [ 65, 128]: nb 8 0.9784 ms, nb 16 0.9722 ms
//   float  m in [129, 256]: nb 8 2.5025 ms, nb 16 2.3172 ms
//   float  m in [257, 512]: nb 8 11.5111 ms, nb 16 7.6585 ms
//   double m in [ 16,  64]: nb 8 0.5061 ms, nb 16 0.6048 ms
//   double m in [ 65, 128]: nb 8 1.1812 ms, nb 16 1.195 ms
//   double m in [129, 256]: nb 8 2.9508 ms, nb 16 2.6575 ms
//   double m in [257, 512]: nb 8 16.1349 ms, nb 16 10.0218 ms
#pragma once

static const int kTunedNb[2][4] = {
    /* float  */ {8, 16, 16, 16},
    /* double */ {8, 8, 16, 16},
};
