// Voxelized-volume painter: GGEMSVolumeCreatorManager's label volume and the
// box / tube / sphere painters (src/geometries/GGEMSVolumeCreatorManager.cc,
// GGEMSBox.cc, GGEMSTube.cc, GGEMSSphere.cc; src/kernels/DrawGGEMS{Box,Tube,
// Sphere}.cl).  One pass over the volume per shape: each voxel centre is
// tested and, inside the shape, takes the label converted to the volume's
// element type.  The pass is a coalesced read-modify-write of the element
// array, so it runs at HBM speed; the volume lives on the device from
// ggh_volume_create to ggh_volume_read.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <new>
#include <string>

#include "../../../include/ggems_engine.h"

struct ggh_volume {
  int device = 0;
  int32_t dims[3] = {0, 0, 0};
  float es[3] = {0.f, 0.f, 0.f};
  int32_t type = 0;
  size_t n = 0;
  void* data = nullptr;
  hipStream_t stream = nullptr;
};

namespace {

constexpr int kDrawBlock = 256;

thread_local std::string g_vol_error;

int vfail(int code, const std::string& msg) {
  g_vol_error = msg;
  return code;
}

#define VHIP(expr)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return vfail(-100 - (int)e_, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

size_t element_bytes(int32_t t) {
  switch (t) {
    case GGH_MET_CHAR:
    case GGH_MET_UCHAR: return 1;
    case GGH_MET_SHORT:
    case GGH_MET_USHORT: return 2;
    case GGH_MET_INT:
    case GGH_MET_UINT:
    case GGH_MET_FLOAT: return 4;
    default: return 0;
  }
}

struct DrawArgs {
  uint64_t n;
  float es[3];
  int32_t dims[3];
  float pos[3];
  float label;
  int32_t shape;
  float a, b, c;  // box: height, width, depth; tube: height, radius_x, radius_y; sphere: radius
};

// voxel indices the way the reference kernels derive them: the y index
// takes the id modulo dims.x * dims.x, which is only right for square x/y
// (DrawGGEMSTube.cl:82, DrawGGEMSBox.cl:87); kept so that every volume is
// painted identically
__device__ inline void voxel_indices(uint64_t gid, const int32_t d[3], int32_t& ix, int32_t& iy, int32_t& iz) {
  const uint64_t xx = (uint64_t)d[0] * (uint64_t)d[0], xy = (uint64_t)d[0] * (uint64_t)d[1];
  iy = (int32_t)((gid % xx) / (uint64_t)d[0]);
  ix = (int32_t)((int64_t)(gid % xy) - (int64_t)iy * d[0]);
  iz = (int32_t)(gid / xy);
}

// centre of voxel i along an axis: (size / 2) * (1 - n + 2 i), minus the shape position
__device__ inline float centre(float es, int32_t n, int32_t i, float pos) {
  const float h = es / 2.0f;
  const float k = (1.0f - (float)n) + 2.0f * (float)i;
  return h * k - pos;
}

__device__ inline bool inside(const DrawArgs& A, float x, float y, float z) {
  if (A.shape == GGH_SHAPE_BOX) {
    // height along y, width along x, depth along z (DrawGGEMSBox.cl:92-96)
    const float hh = A.a / 2.0f, hw = A.b / 2.0f, hd = A.c / 2.0f;
    return (z <= hd && z >= -hd) && (y <= hh && y >= -hh) && (x <= hw && x >= -hw);
  }
  if (A.shape == GGH_SHAPE_TUBE) {
    // along z, elliptic section (DrawGGEMSTube.cl:89-90)
    if (!(z <= A.a / 2.0f && z >= -A.a / 2.0f)) return false;
    const float tx = x * x / (A.b * A.b);
    const float ty = y * y / (A.c * A.c);
    return tx + ty <= 1.0f;
  }
  // sphere (DrawGGEMSSphere.cl:84)
  const float r2 = x * x;
  const float s2 = r2 + y * y;
  const float t2 = s2 + z * z;
  return t2 <= A.a * A.a;
}

template <typename T>
__global__ __launch_bounds__(kDrawBlock) void draw_kernel(DrawArgs A, T* vol) {
  for (uint64_t gid = blockIdx.x * (uint64_t)kDrawBlock + threadIdx.x; gid < A.n;
       gid += (uint64_t)gridDim.x * kDrawBlock) {
    int32_t ix, iy, iz;
    voxel_indices(gid, A.dims, ix, iy, iz);
    const float x = centre(A.es[0], A.dims[0], ix, A.pos[0]);
    const float y = centre(A.es[1], A.dims[1], iy, A.pos[1]);
    const float z = centre(A.es[2], A.dims[2], iz, A.pos[2]);
    if (inside(A, x, y, z)) vol[gid] = (T)A.label;
  }
}

template <typename T>
int launch_draw(const ggh_volume* v, const DrawArgs& A) {
  const uint64_t blocks = (A.n + kDrawBlock - 1) / kDrawBlock;
  const unsigned grid = (unsigned)(blocks < 65536 ? blocks : 65536);
  hipLaunchKernelGGL(draw_kernel<T>, dim3(grid), dim3(kDrawBlock), 0, v->stream, A, static_cast<T*>(v->data));
  VHIP(hipGetLastError());
  VHIP(hipStreamSynchronize(v->stream));
  return 0;
}

}  // namespace

extern "C" {

const char* ggh_volume_last_error(void) { return g_vol_error.c_str(); }

int ggh_volume_create(int device, const int32_t dims[3], const float element_size[3], int32_t data_type,
                      ggh_volume** out) {
  if (!out || !dims || !element_size) return vfail(-1, "null argument");
  *out = nullptr;
  const size_t eb = element_bytes(data_type);
  if (!eb) return vfail(-2, "unknown element type");
  if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return vfail(-2, "volume dimensions have to be > 0");
  int n_dev = 0;
  VHIP(hipGetDeviceCount(&n_dev));
  if (device < 0 || device >= n_dev) return vfail(-2, "device index out of range");
  VHIP(hipSetDevice(device));
  ggh_volume* v = new (std::nothrow) ggh_volume();
  if (!v) return vfail(-3, "out of host memory");
  v->device = device;
  for (int i = 0; i < 3; ++i) {
    v->dims[i] = dims[i];
    v->es[i] = element_size[i];
  }
  v->type = data_type;
  v->n = (size_t)dims[0] * dims[1] * dims[2];
  hipError_t e = hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&v->data, v->n * eb);
  if (e == hipSuccess) e = hipMemsetAsync(v->data, 0, v->n * eb, v->stream);  // CleanBuffer
  if (e == hipSuccess) e = hipStreamSynchronize(v->stream);
  if (e != hipSuccess) {
    ggh_volume_destroy(v);
    return vfail(-100 - (int)e, std::string("volume allocation: ") + hipGetErrorString(e));
  }
  *out = v;
  return 0;
}

int ggh_volume_draw(ggh_volume* v, const ggh_volume_shape* s) {
  if (!v || !s) return vfail(-1, "null argument");
  if (s->shape != GGH_SHAPE_BOX && s->shape != GGH_SHAPE_TUBE && s->shape != GGH_SHAPE_SPHERE)
    return vfail(-2, "unknown shape");
  VHIP(hipSetDevice(v->device));
  DrawArgs A;
  std::memset(&A, 0, sizeof A);
  A.n = v->n;
  for (int i = 0; i < 3; ++i) {
    A.es[i] = v->es[i];
    A.dims[i] = v->dims[i];
    A.pos[i] = s->position[i];
  }
  A.label = s->label;
  A.shape = s->shape;
  A.a = s->size[0];
  A.b = s->size[1];
  A.c = s->size[2];
  switch (v->type) {
    case GGH_MET_CHAR: return launch_draw<int8_t>(v, A);
    case GGH_MET_UCHAR: return launch_draw<uint8_t>(v, A);
    case GGH_MET_SHORT: return launch_draw<int16_t>(v, A);
    case GGH_MET_USHORT: return launch_draw<uint16_t>(v, A);
    case GGH_MET_INT: return launch_draw<int32_t>(v, A);
    case GGH_MET_UINT: return launch_draw<uint32_t>(v, A);
    default: return launch_draw<float>(v, A);
  }
}

int ggh_volume_read(ggh_volume* v, void* host, size_t bytes) {
  if (!v || !host) return vfail(-1, "null argument");
  if (bytes != v->n * element_bytes(v->type)) return vfail(-2, "buffer size does not match the volume");
  VHIP(hipSetDevice(v->device));
  VHIP(hipMemcpyAsync(host, v->data, bytes, hipMemcpyDeviceToHost, v->stream));
  VHIP(hipStreamSynchronize(v->stream));
  return 0;
}

int ggh_volume_destroy(ggh_volume* v) {
  if (!v) return 0;
  (void)hipSetDevice(v->device);
  if (v->stream) (void)hipStreamSynchronize(v->stream);
  if (v->data) (void)hipFree(v->data);
  if (v->stream) (void)hipStreamDestroy(v->stream);
  delete v;
  return 0;
}

}  // extern "C"
