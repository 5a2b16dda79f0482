// sdem_host.h -- entry points that the library's own host modules (the
// surrogate in gnn_api.cpp, the file formats in sdem_io.cpp) use to reach a
// context's internals.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sdem.h"

// device view of a single-box context's current state buffer
struct sdem_dem_view {
  double4* pos;
  double* lv;
  double* la;
  int* id;
  int cap, n;
  double dt, g[3], lo[3], hi[3];
  int per[3];
  hipStream_t stream;
};
extern "C" int sdem_internal_view(sdem_ctx* c, sdem_dem_view* v);
extern "C" int sdem_internal_after_update(sdem_ctx* c, int32_t steps);

// what the file readers and writers need of the selected box; the pointers
// stay valid until the next set-up call on the context
struct sdem_box_info {
  int32_t tsize, rsize;        // levels of linVel/linAcc and rotVel/rotAcc
  int32_t num_prtcl;           // DEM_opt%numPrtcl
  int32_t ntype, nwall, nfix;
  int32_t keep_forces;
  double lo[3], hi[3];
  int32_t periodic[3];
  const double* type_radius;   // ntype radii (Prtcl_PureProp%Radius)
  const int32_t* wall_id;      // nwall wall ids
  const int32_t* fix_id;       // nfix fixed-particle ids
};
extern "C" int sdem_internal_box_info(sdem_ctx* c, sdem_box_info* out);

// GPrtcl_UsrMark of the given particles (by id) set in place, without
// touching the rest of the state or the contact history
extern "C" int sdem_internal_set_marks(sdem_ctx* c, int32_t n, const int32_t* id, const int32_t* mark);

// record the message of a failure on the context (sdem_last_error)
extern "C" int sdem_internal_fail(sdem_ctx* c, int32_t code, const char* msg);
