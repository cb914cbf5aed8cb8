// mmf_output.cpp -- VTK output of a context's state (include/mmf.h).
//
// The reference writes, at every save (main.cpp:284-298, 346, 509-540), the
// cell fields of solver_writer.cpp:56-119 through bitpit's VTK writer after a
// host conservative-to-primitive pass over every cell.  Here the conversion is
// one GPU kernel into a device staging buffer; an asynchronous copy moves it
// to one of two pinned host buffers and a host thread writes the file (VTK XML
// image data, appended raw arrays), so the disk write of one save overlaps the
// next steps.  Two saves can be in flight; a third waits for the older one.
// The writer thread makes no HIP call: a host function enqueued on the stream
// behind the copy marks the buffer ready (a HIP synchronisation from another
// thread would invalidate a graph capture running on the stream meanwhile).
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <thread>

#include "mmf.h"
#include "mmf_ctx.hpp"
#include "mmf_internal.hpp"

struct mmf_output {
  mmf::OutputLayout layout{};
  char *dstage = nullptr;
  char *hbuf[2] = {nullptr, nullptr};
  bool busy[2] = {false, false};   // a write is pending on the buffer
  bool ready[2] = {false, false};  // its copy has completed
  struct Slot {
    mmf_output *o;
    int b;
  } slot[2];
  int next = 0;
  struct Job {
    int buf;
    std::string path;
  };
  std::deque<Job> jobs;
  std::mutex mu;
  std::condition_variable cv;
  bool stop = false;
  std::string error;  // first failed write since the last wait
  std::thread worker;
  GridDev g{};
  double origin[3] = {0.0, 0.0, 0.0};
};

namespace {

int ofail(mmf_ctx *ctx, int code, const std::string &msg) {
  ctx->err = msg;
  return code;
}

// One field of the VTK file: name, VTK type, components, byte offset in the
// staging buffer, byte count.
struct Field {
  const char *name;
  const char *type;
  int comps;
  int64_t offset, bytes;
};

bool write_vti(const std::string &path, const GridDev &g, const double *origin, const char *data,
               const mmf::OutputLayout &L, std::string &err) {
  const int64_t N = g.N;
  const Field fields[] = {
      {"solved", "Int32", 1, L.solved, 4 * N},          {"velocity", "Float64", 3, L.vel, 24 * N},
      {"pressure", "Float64", 1, L.p, 8 * N},           {"temperature", "Float64", 1, L.T, 8 * N},
      {"density", "Float64", 1, L.rho, 8 * N},          {"residualC", "Float64", 1, L.res[0], 8 * N},
      {"residualMX", "Float64", 1, L.res[1], 8 * N},    {"residualMY", "Float64", 1, L.res[2], 8 * N},
      {"residualMZ", "Float64", 1, L.res[3], 8 * N},    {"residualE", "Float64", 1, L.res[4], 8 * N},
  };
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  const long long nx = g.n[0], ny = g.n[1], nz = (g.dim == 3) ? g.n[2] : 1;
  std::fprintf(f, "<?xml version=\"1.0\"?>\n");
  std::fprintf(f, "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n");
  std::fprintf(f, "  <ImageData WholeExtent=\"0 %lld 0 %lld 0 %lld\" Origin=\"%.17g %.17g %.17g\" Spacing=\"%.17g %.17g %.17g\">\n",
               nx, ny, nz, origin[0], origin[1], origin[2], g.h, g.h, g.h);
  std::fprintf(f, "    <Piece Extent=\"0 %lld 0 %lld 0 %lld\">\n", nx, ny, nz);
  std::fprintf(f, "      <CellData Scalars=\"density\" Vectors=\"velocity\">\n");
  long long off = 0;
  for (const Field &fd : fields) {
    std::fprintf(f, "        <DataArray type=\"%s\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"appended\" offset=\"%lld\"/>\n",
                 fd.type, fd.name, fd.comps, off);
    off += 8 + (long long)fd.bytes;
  }
  std::fprintf(f, "      </CellData>\n    </Piece>\n  </ImageData>\n  <AppendedData encoding=\"raw\">\n_");
  bool ok = true;
  for (const Field &fd : fields) {
    const uint64_t nb = (uint64_t)fd.bytes;
    ok = ok && std::fwrite(&nb, sizeof(nb), 1, f) == 1;
    ok = ok && (nb == 0 || std::fwrite(data + fd.offset, 1, nb, f) == nb);
  }
  std::fprintf(f, "\n  </AppendedData>\n</VTKFile>\n");
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) err = "write failed: " + path;
  return ok;
}

// Runs on the runtime's callback thread once the copy into buffer b is done.
void copy_done(void *arg) {
  auto *sl = static_cast<mmf_output::Slot *>(arg);
  {
    std::lock_guard<std::mutex> lk(sl->o->mu);
    sl->o->ready[sl->b] = true;
  }
  sl->o->cv.notify_all();
}

void worker_loop(mmf_output *o) {
  for (;;) {
    mmf_output::Job job;
    {
      std::unique_lock<std::mutex> lk(o->mu);
      o->cv.wait(lk, [o] { return !o->jobs.empty() ? o->ready[o->jobs.front().buf] : o->stop; });
      if (o->jobs.empty()) return;  // stop requested and nothing left
      job = o->jobs.front();
    }
    std::string err;
    const bool ok = write_vti(job.path, o->g, o->origin, o->hbuf[job.buf], o->layout, err);
    {
      std::lock_guard<std::mutex> lk(o->mu);
      if (!ok && o->error.empty()) o->error = err;
      o->jobs.pop_front();
      o->ready[job.buf] = false;
      o->busy[job.buf] = false;
    }
    o->cv.notify_all();
  }
}

int ensure_output(mmf_ctx *ctx) {
  if (ctx->out) return MMF_OK;
  mmf_output *o = new (std::nothrow) mmf_output();
  if (!o) return ofail(ctx, MMF_ERR_OUT_OF_MEMORY, "output state");
  o->layout = mmf::output_layout(ctx->g.N);
  o->g = ctx->g;
  for (int b = 0; b < 2; ++b) o->slot[b] = {o, b};
  for (int a = 0; a < 3; ++a) o->origin[a] = (a < ctx->g.dim) ? ctx->desc.origin[a] : 0.0;
  const size_t bytes = (size_t)o->layout.total;
  bool ok = hipMalloc(&o->dstage, bytes) == hipSuccess;
  for (int b = 0; b < 2 && ok; ++b) ok = hipHostMalloc(&o->hbuf[b], bytes, hipHostMallocDefault) == hipSuccess;
  if (!ok) {
    ctx->out = o;
    mmf_output_release(ctx);
    return ofail(ctx, MMF_ERR_OUT_OF_MEMORY, "output buffers");
  }
  o->worker = std::thread(worker_loop, o);
  ctx->out = o;
  return MMF_OK;
}

}  // namespace

void mmf_output_release(mmf_ctx *ctx) {
  mmf_output *o = ctx->out;
  if (!o) return;
  if (o->worker.joinable()) {
    {
      std::lock_guard<std::mutex> lk(o->mu);
      o->stop = true;
    }
    o->cv.notify_all();
    o->worker.join();
  }
  for (int b = 0; b < 2; ++b)
    if (o->hbuf[b]) (void)hipHostFree(o->hbuf[b]);
  if (o->dstage) (void)hipFree(o->dstage);
  delete o;
  ctx->out = nullptr;
}

extern "C" {

int mmf_write_vtk_async(mmf_ctx *ctx, const char *path) {
  if (!ctx || !path || !*path) return MMF_ERR_INVALID;
  int rc = ensure_output(ctx);
  if (rc != MMF_OK) return rc;
  mmf_output *o = ctx->out;
  int b;
  {
    std::unique_lock<std::mutex> lk(o->mu);
    b = o->next;
    o->cv.wait(lk, [o, b] { return !o->busy[b]; });
    o->busy[b] = true;
    o->next ^= 1;
  }
  auto release = [&](int code, const std::string &msg) {
    {
      std::lock_guard<std::mutex> lk(o->mu);
      o->busy[b] = false;
    }
    o->cv.notify_all();
    return ofail(ctx, code, msg);
  };
  // residual fields: the RHS of the state being written
  rc = mmf_ctx_enqueue_rhs(ctx);
  if (rc != MMF_OK) return release(rc, ctx->err);
  hipError_t e = mmf::launch_output_fields(ctx->g, ctx->dU, ctx->dR, ctx->g.flags, o->dstage, o->layout, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(o->hbuf[b], o->dstage, (size_t)o->layout.total, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipLaunchHostFunc(ctx->stream, copy_done, &o->slot[b]);
  if (e != hipSuccess) return release(MMF_ERR_HIP, std::string("output: ") + hipGetErrorString(e));
  {
    std::lock_guard<std::mutex> lk(o->mu);
    o->jobs.push_back({b, std::string(path)});
  }
  o->cv.notify_all();
  return MMF_OK;
}

int mmf_wait_writes(mmf_ctx *ctx) {
  if (!ctx) return MMF_ERR_INVALID;
  mmf_output *o = ctx->out;
  if (!o) return MMF_OK;
  std::string err;
  {
    std::unique_lock<std::mutex> lk(o->mu);
    o->cv.wait(lk, [o] { return o->jobs.empty(); });
    err.swap(o->error);
  }
  if (!err.empty()) return ofail(ctx, MMF_ERR_STATE, err);
  return MMF_OK;
}

int mmf_write_vtk(mmf_ctx *ctx, const char *path) {
  int rc = mmf_write_vtk_async(ctx, path);
  if (rc != MMF_OK) return rc;
  return mmf_wait_writes(ctx);
}

}  // extern "C"
