// Serving loop (placeholder until the request loop lands).
#include "a3d_handles.h"

using namespace a3d;

extern "C" {
int a3d_server_create(const a3d_server_config*, a3d_server**) { return fail(A3D_ERR_STATE, "serving: not built"); }
int a3d_server_destroy(a3d_server*) { return A3D_OK; }
int a3d_server_add_map(a3d_server*, const a3d_grid*, int*) { return fail(A3D_ERR_STATE, "serving: not built"); }
int a3d_server_open_session(a3d_server*, int, int, const float*, const float*, uint64_t, int*) {
  return fail(A3D_ERR_STATE, "serving: not built");
}
int a3d_server_close_session(a3d_server*, int) { return fail(A3D_ERR_STATE, "serving: not built"); }
int a3d_server_submit(a3d_server*, int, const double*, const double*, const float*, const float*, const float*,
                      size_t, const float*, size_t, double, double, double, double, int, uint64_t*) {
  return fail(A3D_ERR_STATE, "serving: not built");
}
int a3d_server_poll(a3d_server*, a3d_result*, size_t, size_t*) { return fail(A3D_ERR_STATE, "serving: not built"); }
int a3d_server_drain(a3d_server*) { return fail(A3D_ERR_STATE, "serving: not built"); }
int a3d_server_get_particles(a3d_server*, int, float*, size_t) { return fail(A3D_ERR_STATE, "serving: not built"); }
int a3d_server_stats(a3d_server*, uint64_t*, uint64_t*, uint64_t*, double*) {
  return fail(A3D_ERR_STATE, "serving: not built");
}
}
