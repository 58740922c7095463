// Generated by mat_amd/autotune_batched.py on one MI355X: panel width of
// the batched uneven QR per m bin, the faster of 8 and 16 on 1500 matrices
// per bin (n uniform in [8, m]), median of 3 hipGraph replays.
//
//   float  m in [ 16,  64]: nb 8 0.2067 ms, nb 16 0.1705 ms
//   float  m in [ 65, 128]: nb 8 0.4963 ms, nb 16 0.329 ms
//   float  m in [129, 256]: nb 8 1.6862 ms, nb 16 1.0581 ms
//   float  m in [257, 512]: nb 8 9.6493 ms, nb 16 5.4351 ms
//   double m in [ 16,  64]: nb 8 0.32 ms, nb 16 0.244 ms
//   double m in [ 65, 128]: nb 8 0.7412 ms, nb 16 0.5783 ms
//   double m in [129, 256]: nb 8 2.3659 ms, nb 16 1.4707 ms
//   double m in [257, 512]: nb 8 14.9863 ms, nb 16 8.2751 ms
#pragma once

static const int kTunedNb[2][4] = {
    /* float  */ {16, 16, 16, 16},
    /* double */ {16, 16, 16, 16},
};
