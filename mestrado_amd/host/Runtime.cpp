#include "Runtime.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace mstrt {
namespace {
bool g_have_subst = false;
int8_t g_subst[576];
int g_device = 0, g_form = MST_FORM_BATCHED;
mst_ctx* g_ctx = nullptr;
int g_go = -1, g_ge = -1;
}  // namespace

void configure(const int8_t* subst, int device, int f) {
  if (subst) {
    memcpy(g_subst, subst, 576);
    g_have_subst = true;
  }
  g_device = device;
  g_form = f;
}

mst_ctx* context(int go, int ge) {
  if (g_ctx && go == g_go && ge == g_ge) return g_ctx;
  if (g_ctx) {
    mst_destroy(g_ctx);
    g_ctx = nullptr;
  }
  int rc = mst_create(go, ge, g_have_subst ? g_subst : nullptr, g_device, &g_ctx);
  if (rc != MST_OK) {
    fprintf(stderr, "There are no available device(s) or the settings are invalid (mst_create %d)\n", rc);
    exit(rc == MST_E_ARG ? 1 : 0);
  }
  g_go = go;
  g_ge = ge;
  return g_ctx;
}

int form() { return g_form; }

void check(int rc, const char* what) {
  if (rc == MST_OK) return;
  fprintf(stderr, "GPUassert: %s failed (%d): %s\n", what, rc, g_ctx ? mst_last_error(g_ctx) : "");
  exit(rc < 0 ? -rc : rc);
}

void shutdown() {
  if (g_ctx) mst_destroy(g_ctx);
  g_ctx = nullptr;
}

}  // namespace mstrt
