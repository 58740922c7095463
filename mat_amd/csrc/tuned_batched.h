// Generated by mat_amd/autotune_batched.py on one MI355X.
// Panel width of the batched uneven QR per m bin: the faster of 8 and 16 on 1500
// matrices per bin (n uniform in [8, m]); residency and queue order: the fastest
// pair on 5000 matrices with m ~ U{16..512}, n ~ U{8..m}; median of 3 hipGraph
// replays each.
//
//   float  m in [ 16,  64]: nb 8 0.1955 ms, nb 16 0.1448 ms
//   float  m in [ 65, 128]: nb 8 0.481 ms, nb 16 0.3075 ms
//   float  m in [129, 256]: nb 8 1.5056 ms, nb 16 0.8574 ms
//   float  m in [257, 512]: nb 8 7.2245 ms, nb 16 3.8806 ms
//   float  schedules (per CU / order: ms): 1/0 9.4761, 1/1 11.4762, 1/2 9.6494, 2/0 5.8345, 2/1 7.6313, 2/2 6.0103, 3/0 4.8843, 3/1 6.7469, 3/2 5.0424
//   double m in [ 16,  64]: nb 8 0.299 ms, nb 16 0.2219 ms
//   double m in [ 65, 128]: nb 8 0.6951 ms, nb 16 0.463 ms
//   double m in [129, 256]: nb 8 1.9629 ms, nb 16 1.1754 ms
//   double m in [257, 512]: nb 8 9.6876 ms, nb 16 5.2652 ms
//   double schedules (per CU / order: ms): 1/0 12.2225, 1/1 14.5275, 1/2 12.4592, 2/0 8.6778, 2/1 11.0504, 2/2 8.7993
#pragma once

static const int kTunedNb[2][4] = {
    /* float  */ {16, 16, 16, 16},
    /* double */ {16, 16, 16, 16},
};

// Schedule: resident workgroups per CU (0: the occupancy limit) and queue
// order (0: decreasing flop count, 1: as given, 2: largest and smallest
// alternating), [float, double].
static const int kTunedPerCu[2] = {3, 2};
static const int kTunedOrder[2] = {0, 0};
