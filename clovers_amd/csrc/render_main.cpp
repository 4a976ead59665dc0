/*
 * clovers-render -- the reference's `clovers render` subcommand
 * (clovers-cli/src/render.rs:17-182) over the C ABI: read a scene file,
 * render it on the GPU, write PNG and/or EXR.
 *
 *   clovers-render -i scenes/cornell.json -w 1024 -h 1024 -s 64 -d 64 \
 *                  -f png exr -o renders/cornell [--bvh sah|lam] [--seed N] [--guiding]
 *
 * Options and defaults are RenderOptions' (render.rs:17-49), with the modes
 * path-tracing, normal-map, bvh-test-count and primitive-test-count
 * (render.rs:52-62); `random` is the only sampler (blue is reported as
 * unsupported).  The output path has no extension, as in the reference; the
 * default is renders/<unix time>.  Mesh paths in the scene file are relative
 * to the working directory, or to --mesh-root.  --guiding and --builder
 * (the scene BVH built on the host or on the GPU) have no reference
 * counterpart.
 */
#include "../../include/clovers_hip.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <string>
#include <vector>

namespace {

int usage(const char *msg) {
  if (msg) fprintf(stderr, "error: %s\n", msg);
  fprintf(stderr,
          "usage: clovers-render -i <scene.json> [-o <output without extension>] [-w 1024] [-h 1024]\n"
          "                      [-s 64] [-d 64] [-m path-tracing|normal-map|bvh-test-count|primitive-test-count]\n"
          "                      [--sampler random] [--bvh sah|lam]\n"
          "                      [-f png [exr]] [--seed N] [--guiding] [--mesh-root DIR] [--device N]\n"
          "                      [--builder host|device]\n");
  return 2;
}

bool read_text(const std::string &path, std::string *out) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return false;
  char buf[1 << 16];
  size_t n;
  while ((n = fread(buf, 1, sizeof(buf), f)) > 0) out->append(buf, n);
  fclose(f);
  return true;
}

}  // namespace

int main(int argc, char **argv) {
  std::string input, output, mode = "path-tracing", sampler = "random", bvh = "sah", mesh_root, builder = "host";
  std::vector<std::string> formats;
  uint32_t width = 1024, height = 1024, samples = 64, max_depth = 64, guiding = 0;
  uint64_t seed = 1;
  int device = 0;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    auto val = [&]() -> const char * {
      if (i + 1 >= argc) { usage(("missing value for " + a).c_str()); exit(2); }
      return argv[++i];
    };
    auto u32 = [&]() {
      const char *v = val();
      char *e;
      unsigned long x = strtoul(v, &e, 10);
      if (*e || x > 0xFFFFFFFFul) { usage(("bad number for " + a).c_str()); exit(2); }
      return (uint32_t)x;
    };
    if (a == "-i" || a == "--input") input = val();
    else if (a == "-o" || a == "--output") output = val();
    else if (a == "-w" || a == "--width") width = u32();
    else if (a == "-h" || a == "--height") height = u32();
    else if (a == "-s" || a == "--samples") samples = u32();
    else if (a == "-d" || a == "--max-depth") max_depth = u32();
    else if (a == "-m" || a == "--mode") mode = val();
    else if (a == "--sampler") sampler = val();
    else if (a == "--bvh") bvh = val();
    else if (a == "--seed") seed = strtoull(val(), nullptr, 10);
    else if (a == "--guiding") guiding = 1;
    else if (a == "--mesh-root") mesh_root = val();
    else if (a == "--device") device = (int)u32();
    else if (a == "--builder") builder = val();
    else if (a == "-f" || a == "--formats") {
      formats.push_back(val());
      while (i + 1 < argc && argv[i + 1][0] != '-') formats.push_back(argv[++i]);
    } else if (a == "--help") return usage(nullptr);
    else return usage(("unknown option " + a).c_str());
  }
  if (input.empty()) return usage("--input is required");
  uint32_t mode_id;
  if (mode == "path-tracing") mode_id = CLV_MODE_PATH_TRACING;
  else if (mode == "normal-map") mode_id = CLV_MODE_NORMAL_MAP;
  else if (mode == "bvh-test-count") mode_id = CLV_MODE_BVH_TEST_COUNT;
  else if (mode == "primitive-test-count") mode_id = CLV_MODE_PRIMITIVE_TEST_COUNT;
  else return usage("unknown mode");
  if (sampler != "random") return usage("only the random sampler is supported");
  if (bvh != "sah" && bvh != "lam") return usage("--bvh is sah or lam");
  if (builder != "host" && builder != "device") return usage("--builder is host or device");
  if (formats.empty()) formats.push_back("png");
  for (const std::string &f : formats)
    if (f != "png" && f != "exr") return usage("formats are png and exr");
  if (output.empty()) output = "renders/" + std::to_string((long long)time(nullptr));

  std::string json;
  if (!read_text(input, &json)) {
    fprintf(stderr, "error: cannot read %s\n", input.c_str());
    return 1;
  }
  char err[512];
  clv_scene_file *sf = nullptr;
  clv_status st = clv_scene_file_parse(json.c_str(), mesh_root.empty() ? nullptr : mesh_root.c_str(),
                                       bvh == "sah" ? CLV_BVH_SAH : CLV_BVH_LAM, &sf, err, sizeof(err));
  if (st != CLV_OK) {
    fprintf(stderr, "error: %s: %s\n", input.c_str(), err);
    return 1;
  }
  clv_ctx *ctx = nullptr;
  if ((st = clv_create(device, &ctx)) != CLV_OK) {
    fprintf(stderr, "error: no HIP device %d (status %d)\n", device, (int)st);
    clv_scene_file_free(sf);
    return 1;
  }
  uint32_t sid = 0;
  if ((st = clv_scene_upload_ex(ctx, clv_scene_file_desc(sf), builder == "device" ? CLV_BUILD_DEVICE : CLV_BUILD_HOST,
                                &sid)) != CLV_OK) {
    fprintf(stderr, "error: scene upload: %s\n", clv_last_error(ctx));
    clv_destroy(ctx);
    clv_scene_file_free(sf);
    return 1;
  }
  clv_scene_file_free(sf);
  clv_render_opts o;
  memset(&o, 0, sizeof(o));
  o.width = width;
  o.height = height;
  o.samples = samples;
  o.max_depth = max_depth;
  o.seed = seed;
  o.guiding = guiding;
  std::vector<float> img((size_t)width * height * 3);
  printf("Rendering %s at %ux%u, %u samples per pixel, max depth %u\n", input.c_str(), width, height, samples,
         max_depth);
  auto t0 = std::chrono::steady_clock::now();
  st = mode_id == CLV_MODE_PATH_TRACING ? clv_render(ctx, sid, &o, img.data(), nullptr)
                                       : clv_render_debug(ctx, sid, &o, mode_id, img.data());
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (st != CLV_OK) {
    fprintf(stderr, "error: render: %s\n", clv_last_error(ctx));
    clv_destroy(ctx);
    return 1;
  }
  clv_destroy(ctx);
  const double msps = (double)width * height * samples / secs / 1e6;
  printf("Rendering finished in %.3f s (%.1f Msamples/s)\n", secs, msps);
  char comment[1024];
  snprintf(comment, sizeof(comment),
           "Rendered with clovers_amd. Scene file %s rendered using the %s rendering mode at %ux%u resolution%s "
           "Rendering finished in %.3f s on one GPU.",
           input.c_str(), mode.c_str(), width, height,
           mode_id == CLV_MODE_PATH_TRACING
               ? (", " + std::to_string(samples) + " samples per pixel, " + std::to_string(max_depth) +
                  " max ray bounce depth.").c_str()
               : ".",
           secs);
  {
    std::error_code ec;
    std::filesystem::path parent = std::filesystem::path(output).parent_path();
    if (!parent.empty()) std::filesystem::create_directories(parent, ec);
  }
  for (const std::string &f : formats) {
    std::string path = output + "." + f;
    st = f == "png" ? clv_write_png(img.data(), width, height, path.c_str(), comment)
                    : clv_write_exr(img.data(), width, height, path.c_str());
    if (st != CLV_OK) {
      fprintf(stderr, "error: cannot write %s\n", path.c_str());
      return 1;
    }
    printf("Wrote %s\n", path.c_str());
  }
  return 0;
}
