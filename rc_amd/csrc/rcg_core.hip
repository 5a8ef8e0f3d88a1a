// rcg_core.hip -- error state and argument checks shared by all entry points.
#include <cstdarg>
#include <cstring>

#include "rcg_internal.h"

namespace rcg {

namespace {
thread_local int t_err = 0;
thread_local char t_msg[512] = {0};
}  // namespace

void set_error(int code, const char *fmt, ...)
{
  t_err = code;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_msg, sizeof(t_msg), fmt, ap);
  va_end(ap);
}

void clear_error()
{
  t_err = 0;
  t_msg[0] = 0;
}

rcg_status check_geom(const rcg_geom *g)
{
  if (g->n_chunks < 0 || g->ts_ch < 0) {
    set_error(RCG_ERR_ARG, "negative chunk count (%d) or chunk length (%d)",
              g->n_chunks, g->ts_ch);
    return RCG_ERR_ARG;
  }
  if (g->grid_size <= 0 || g->grid_size > 32767 || g->grid_pitch < g->grid_size) {
    set_error(RCG_ERR_ARG,
              "grid_size must be in [1, 32767] and grid_pitch >= grid_size "
              "(got size %d, pitch %d)", g->grid_size, g->grid_pitch);
    return RCG_ERR_ARG;
  }
  if ((long long)g->grid_size * g->grid_pitch > 0x7fffffffLL) {
    set_error(RCG_ERR_ARG, "grid of %d x %d cells exceeds 2^31 cells",
              g->grid_size, g->grid_pitch);
    return RCG_ERR_ARG;
  }
  if ((long long)g->n_chunks * g->ts_ch > 0x7fffffffLL) {
    set_error(RCG_ERR_ARG, "%lld points exceed 2^31 - 1",
              (long long)g->n_chunks * g->ts_ch);
    return RCG_ERR_ARG;
  }
  if ((long long)g->n_chunks * g->ts_ch > 0 && (!g->uvw || !g->chunk_supps)) {
    set_error(RCG_ERR_ARG, "null uvw or chunk support array");
    return RCG_ERR_ARG;
  }
  return RCG_OK;
}

rcg_status check_gcf(const rcg_gcf *t)
{
  if (!t->data || !t->layer_off || !t->layer_len || !t->supps) {
    set_error(RCG_ERR_ARG, "null GCF table array");
    return RCG_ERR_ARG;
  }
  if (t->n_slots <= 0 || t->max_w_plane < 0) {
    set_error(RCG_ERR_ARG, "GCF table: n_slots %d, max_w_plane %d",
              t->n_slots, t->max_w_plane);
    return RCG_ERR_ARG;
  }
  if (t->half && t->center != 0) {
    set_error(RCG_ERR_ARG, "half GCF table must have center 0 (got %d)",
              t->center);
    return RCG_ERR_ARG;
  }
  return RCG_OK;
}

}  // namespace rcg

extern "C" int rcg_last_error(void) { return rcg::t_err; }
extern "C" const char *rcg_last_error_string(void) { return rcg::t_msg; }
