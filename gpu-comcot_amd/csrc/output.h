// Reference-format output files of the device loop.
//
// The reference's main loop writes, on the host and synchronously:
//   ALL_PRT / DATA_PRT (output.f90:737-851) at K = 0 and every IPRT root
//     steps (comcot.f90:308-310, 415-417): z_LL_KKKKKK.dat with dry land
//     shown as 0 and the tide correction, m_/n_LL_KKKKKK.dat when the
//     layer's FLUXSWITCH is 0; one Fortran unformatted sequential record
//     of the whole (NX,NY) plane each;
//   MAX_AMP (output.f90:855-1031) after every root step when OUT_OPT is 0
//     or 2 (comcot.f90:420-423): zmax_layer01_MMMMmins.dat at each half
//     hour of model time and zmax_layer01.dat near the end, from the
//     MAX_AMP maximum of the outermost layer, masked like DATA_PRT.
//
// Here the masking runs on the device into a staging area (one pass per
// layer), the staging area goes to pinned host memory on a copy stream of
// its own, a host function on that stream hands the files to a pool of
// writer threads as soon as the copy ends, and the writers put the records
// on disk in parallel, so the step loop only pays for the pack pass.  Three
// staging slots are used in turn; a fourth output waits for the oldest (the
// loop then runs at the speed of the disk).  A queued write that a later
// output of the same file supersedes is dropped.
#pragma once

#include <hip/hip_runtime.h>

#include "tree.h"

namespace gc {

// dir: directory (created by the caller); iprt <= 0: no snapshots;
// out_opt 0/2: maximum files (the outermost layer's MAX_AMP maximum is then
// kept by the step kernel); tend: the run's TEND; tide: TIDE_LEVEL.
int out_configure(Tree &T, const char *dir, int dirlen, int iprt, int out_opt, float tend,
                  float tide);
// FLUXSWITCH of layer lid: 0 z, m, n; 1 or 9 z; others nothing (default 1).
int out_set_fluxswitch(int lid, int fluxswitch);
// Stop writing: waits for queued files, then forgets the configuration
// and the file count.
int out_disable(Tree &T);
bool out_active();

// Root steps from `done` completed ones until the next root step that
// writes files, at most n (n when none comes within n).
long long out_steps_to_next(const Tree &T, long long done, long long n);
// Whether root step k writes files (k = 0: the initial state).
bool out_due(const Tree &T, long long k);
// Queue the files of root step k on stream st (the step's stream).
int out_write(Tree &T, long long k, hipStream_t st);
// Wait until every queued file is on disk; *nfiles = files written since
// out_configure.  Returns the first write error, if any.
int out_flush(long long *nfiles);

}  // namespace gc
