// Kernel shape parameters of libcpmgpu.so, chosen by cpm-ibc_amd/tune.py on one
// MI355X (gfx950); a build can override them with -D.
//
//   CPMGPU_XCD_RUN   consecutive row blocks each XCD takes (workgroups are
//                    dispatched to the 8 XCDs round-robin; runs keep a tile's
//                    gathered x rows in that XCD's L2)
//   CPMGPU_LDS_ROWS  longest LDS tile list of the E gather (FP32 x 8 rows of
//                    32 B: the LDS of one workgroup is 32 B x this)
//
// Torus-knot system (405,928 rows): E gather FP32 x 8 per launch and 32
// right-hand sides streamed through 8 FP32+refine lanes:
//   xcd_run  lds_rows  gather (us)  stream (ms)
//         16      2560        36.5        345.3
//         32      2560        37.3        346.7
//         64      2560        36.5        345.3   <- chosen
//        128      2560        35.8        344.8
//        256      2560        37.6        354.9
//         64      3072        55.6        426.2
#pragma once
#ifndef CPMGPU_XCD_RUN
#define CPMGPU_XCD_RUN 64
#endif
#ifndef CPMGPU_LDS_ROWS
#define CPMGPU_LDS_ROWS 2560
#endif
