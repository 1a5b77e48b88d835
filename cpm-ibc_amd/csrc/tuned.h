// Kernel shape parameters of libcpmgpu.so, chosen by cpm-ibc_amd/tune.py on one
// MI355X (gfx950); a build can override them with -D.
//
//   CPMGPU_XCD_RUN   consecutive row blocks each XCD takes (workgroups are
//                    dispatched to the 8 XCDs round-robin; runs keep a tile's
//                    gathered x rows in that XCD's L2)
//   CPMGPU_LDS_ROWS  longest list of a row tile of the E gather (FP32 x 8
//                    rows of 32 B: the LDS of one workgroup is 32 B x this)
//   CPMGPU_GROUP_ROWS      rows of one shared-stencil item (one thread)
//   CPMGPU_GROUP_LDS_ROWS  longest list of a shared-stencil tile
//
// Torus-knot system (405,928 rows): E gather FP32 x 8 per launch and 32
// right-hand sides streamed through 8 FP32+refine lanes:
//   xcd_run  lds_rows  grp_rows  grp_lds  gather (us)  stream (ms)
//         64      2560         2     1280        26.1        299.7   <- chosen
//         32      2560         2     1280        26.9        301.4
//        128      2560         2     1280        25.6        298.6
//         64      3072         2     1280        26.2        301.1
//         64      2560         1     1280        26.9        304.1
//         64      2560         3     1280        32.9        324.5
//         64      2560         2     1024        37.7        355.0
//         64      2560         2     1536        30.8        318.5
//
// Measured again once the E gathers took a launch's segments round-robin
// (bench.py's default run: 128 right-hand sides streamed through 64 lanes,
// FP64 batches and one-right-hand-side solves), xcd_run 64 -> 16: streamed
// 161.3 -> 162.5 solves/s, FP64 batches of 64 96.1 -> 96.5, p50 of one
// right-hand side 28.7 -> 28.6 ms, one-RHS FP64 30.5 -> 30.3 solves/s;
// xcd_run 16 chosen (runs of 1/8/32 gave 791.0/795.3/788.7 ms per streamed
// step against 786.8 for 16).
#pragma once
#ifndef CPMGPU_XCD_RUN
#define CPMGPU_XCD_RUN 16
#endif
#ifndef CPMGPU_LDS_ROWS
#define CPMGPU_LDS_ROWS 2560
#endif
#ifndef CPMGPU_GROUP_ROWS
#define CPMGPU_GROUP_ROWS 2
#endif
#ifndef CPMGPU_GROUP_LDS_ROWS
#define CPMGPU_GROUP_LDS_ROWS 1280
#endif
