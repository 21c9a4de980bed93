// The reference's hot-start text snapshots (HOT_START / HOTSTART_DATA,
// hotstart.f90:3-166) to and from a layer tree's device state.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "tree.h"

namespace gc {

// Files per layer and start step: z1, m1, n1, zmax1, zmin1.
constexpr int kHotstartFiles = 5;
// hotstart_read_plane's return value for a file that cannot be opened.
constexpr int kHotstartMissing = -1;

// FORMAT('z1_',I2.2,'_',I6.6,'.dat') and the like: kind 0..4 = z1, m1, n1,
// zmax1, zmin1.
std::string hotstart_name(int kind, int layer_id, int start_step);

// One snapshot file of an nx x ny column-major plane: cells I, J >= first
// (1 on the outermost layer, 2 on nested ones), one '(10F12.5)' WRITE per
// row J.  The reader fills only those cells and returns kHotstartMissing
// when the file cannot be opened.
int hotstart_write_plane(const char *path, const float *plane, int nx, int ny, int first);
int hotstart_read_plane(const char *path, float *plane, int nx, int ny, int first);

// Every layer of T into dir under start step S: level-1 Z, M, N, the
// maximum the device keeps without a wet test (CUDA_GETZMAX's), and
// zeros for Z_MIN.
int tree_hotstart_write(Tree &T, hipStream_t st, int start_step, const std::string &dir);
// START_TYPE 1 (the outermost layer) or 2 (every layer) from dir: every
// file is read and checked before any plane changes; a missing file
// zeroes its whole plane (HOTSTART_DATA's fallback), counted in *missing.
int tree_hotstart_read(Tree &T, hipStream_t st, int start_type, int start_step,
                       const std::string &dir, int *missing);

}  // namespace gc
