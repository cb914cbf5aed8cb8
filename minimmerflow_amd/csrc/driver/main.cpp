// minimmerflow -- command-line front-end of the MI355X time step.
//
// Mirrors the reference executable (main.cpp:46-574): reads ./settings.xml
// with the same keys and defaults (problem.cpp:37-211, main.cpp:97-122,
// body.cpp:49-72), argv[1] overrides discretization.space.nCells, runs the
// SSP-RK3 loop on the GPU through the C-ABI and prints the error line the
// reference's test driver parses (main.cpp:569-573, test_driver.py:42-49).
// argv[2] is nSaves (main.cpp:119-121; default: a save after every step): the
// initial state, every step whose time passes the next save time and the
// final state are written as VTK image data (background_<n>.<counter>.vti,
// final_background_<n>.vti; main.cpp:176-177, 346, 376, 509-540), by a host
// thread while the steps go on.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "mmf.h"
#include "settings.hpp"

namespace {

int problem_type(const std::string &key) {  // problem.cpp:37-59
  if (key == "vortex_xy") return MMF_PROBLEM_VORTEX_XY;
  if (key == "vortex_zx") return MMF_PROBLEM_VORTEX_ZX;
  if (key == "vortex_yz") return MMF_PROBLEM_VORTEX_YZ;
  if (key == "radsod") return MMF_PROBLEM_RADSOD;
  if (key == "sod3d_x") return MMF_PROBLEM_SOD_X;
  if (key == "sod3d_y") return MMF_PROBLEM_SOD_Y;
  if (key == "sod3d_z") return MMF_PROBLEM_SOD_Z;
  if (key == "ffstep") return MMF_PROBLEM_FFSTEP;
  throw std::runtime_error("Problem " + key + " is not supported.");
}

void check(int rc, mmf_ctx *ctx, const char *what) {
  if (rc != MMF_OK) {
    std::string m = std::string(what) + ": " + mmf_status_string(rc);
    if (ctx) m += std::string(" (") + mmf_last_error(ctx) + ")";
    throw std::runtime_error(m);
  }
}

int run(int argc, char **argv) {
  auto root = mmfdrv::read_settings("settings.xml");
  const mmfdrv::Node &cfg = *root;

  const int type = problem_type(cfg.str("problem.type"));
  int dim = (type == MMF_PROBLEM_VORTEX_ZX || type == MMF_PROBLEM_VORTEX_YZ) ? 3
                                                                              : (int)cfg.num("problem.dimensions", 2);
  mmf_problem pb;
  check(mmf_problem_defaults(&pb, type, dim), nullptr, "problem");

  double origin[3] = {pb.domain_origin[0], pb.domain_origin[1], pb.domain_origin[2]};
  if (cfg.has("problem.domain.origin")) {  // problem.cpp:87-95
    origin[0] = cfg.num("problem.domain.origin.x", 0.);
    origin[1] = cfg.num("problem.domain.origin.y", 0.);
    origin[2] = cfg.num("problem.domain.origin.z", 0.);
  }
  double length = pb.domain_length;
  if (cfg.has("problem.domain.length")) length = cfg.num("problem.domain.length");
  else if (!(length > 0.0)) throw std::runtime_error("The lenght of the domain is manditory.");  // problem.cpp:145
  for (int d = 0; d < 3; ++d) pb.domain_origin[d] = origin[d];
  pb.domain_length = length;

  const double tMin = (double)(int)cfg.num("problem.start", 0.);  // read as int (problem.cpp:164)
  double tMax = 0.0;
  if (cfg.has("problem.time.end")) tMax = cfg.num("problem.time.end");
  else check(mmf_problem_default_end_time(&pb, &tMax), nullptr, "end time");

  const int order = (int)cfg.num("discretization.space.order");
  const double cfl = cfg.num("discretization.time.CFL");
  long nCells = (argc > 1) ? std::atol(argv[1]) : (long)cfg.num("discretization.space.nCells");

  std::printf("Domain info:\n  Origin .... %g %g %g\n  Length .... %g\n", origin[0], origin[1], origin[2], length);
  std::printf("Time info:\n  Initial .... %g\n  Final .... %g\n", tMin, tMax);
  std::printf("Space distretization info...\n  Order .... %d\n  Cells per direction .... %ld\n", order, nCells);
  std::printf("Time distretization info:\n  CFL .... %g\n", cfl);

  // bodies (body.cpp:49-72): inclusive boxes
  std::vector<double> boxes;
  if (const mmfdrv::Node *b = cfg.find("bodies")) {
    for (const auto &e : b->children) {
      const std::string shape = e->str("type", "");
      if (shape != "box") throw std::runtime_error("Body shape \"" + shape + "\" is not supported.");
      boxes.insert(boxes.end(), {e->num("xMin"), e->num("yMin"), e->num("zMin"), e->num("xMax"), e->num("yMax"),
                                 e->num("zMax")});
    }
  }

  mmf_grid_desc d{};
  d.dim = dim;
  d.order = order;
  for (int a = 0; a < 3; ++a) {
    d.n[a] = (a < dim) ? nCells : 1;
    d.origin[a] = origin[a];
  }
  d.h = length / (double)nCells;  // VolOctree(dim, origin, length, length / nCells), main.cpp:149
  check(mmf_problem_side_bcs(&pb, &d), nullptr, "border conditions");

  mmf_ctx *ctx = nullptr;
  int rc = mmf_create(&d, &ctx);
  if (rc == MMF_ERR_UNSUPPORTED_ORDER) return 2;  // reconstruction.cpp:75-76: exit(2)
  check(rc, ctx, "create");
  {
    // main.cpp:251-260: every border interface's condition at its centroid
    int64_t nb = 0;
    check(mmf_border_face_count(&d, &nb), ctx, "border conditions");
    std::vector<int32_t> fbc((size_t)nb);
    std::vector<double> finfo((size_t)(5 * nb));
    check(mmf_problem_border_bcs(&pb, &d, fbc.data(), finfo.data()), ctx, "border conditions");
    check(mmf_set_border_bcs(ctx, fbc.data(), finfo.data()), ctx, "border conditions");
  }
  if (!boxes.empty()) check(mmf_set_bodies(ctx, boxes.data(), (int32_t)(boxes.size() / 6), nullptr, 0), ctx, "bodies");
  check(mmf_set_initial_state(ctx, &pb), ctx, "initial conditions");
  check(mmf_set_time(ctx, tMin, tMax, cfl), ctx, "time");

  const int nSaves = (argc > 2) ? std::atoi(argv[2]) : std::numeric_limits<int>::max();
  const std::string base = "background_" + std::to_string(nCells);
  int counter = 0;
  auto save_numbered = [&]() {
    char suffix[32];
    std::snprintf(suffix, sizeof(suffix), ".%04d.vti", counter++);
    check(mmf_write_vtk_async(ctx, (base + suffix).c_str()), ctx, "write");
  };
  save_numbered();  // main.cpp:346

  std::printf("\nStarting computation...\n");
  auto t0 = std::chrono::steady_clock::now();
  int64_t steps = 0;
  double t = tMin;
  double nextSave = tMin;
  while (t < tMax) {
    if (!(nextSave < tMax)) {  // no save can come: the rest in one run
      int64_t k = 0;
      check(mmf_run(ctx, -1, &k, &t), ctx, "run");
      steps += k;
      break;
    }
    check(mmf_advance(ctx, 1), ctx, "advance");
    int64_t st = 0;
    check(mmf_get_time(ctx, &t, nullptr, &st, nullptr), ctx, "time");
    steps += 1;
    if (t > nextSave) {  // main.cpp:509-523
      save_numbered();
      nextSave += (tMax - tMin) / nSaves;
    }
  }
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  check(mmf_write_vtk_async(ctx, ("final_" + base + ".vti").c_str()), ctx, "write");  // main.cpp:528-540
  auto t1 = std::chrono::steady_clock::now();
  check(mmf_wait_writes(ctx), ctx, "write");
  const double disk = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
  std::printf("Steps .... %lld, final time %.17g\n", (long long)steps, t);
  std::printf("Computation time (without disk saving time) is %g\n", secs);
  std::printf("Disk time (waiting for the pending writes) %g\n", disk);

  double l1 = 0.0, l2 = 0.0;
  check(mmf_error_vs_exact(ctx, &pb, tMax, &l1, &l2), ctx, "error");
  int64_t np = 0, nr = 0;
  check(mmf_get_positivity_counters(ctx, &np, &nr), ctx, "counters");
  if (np || nr) std::printf("***** Negative pressure in %lld, negative density in %lld flux evaluations\n",
                            (long long)np, (long long)nr);
  std::printf("\n ::::::::: Error check :::::::::\n\n");
  std::printf(" Final error:  %.12e\n\n", l1);
  mmf_destroy(ctx);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "minimmerflow: %s\n", e.what());
    return 1;
  }
}
