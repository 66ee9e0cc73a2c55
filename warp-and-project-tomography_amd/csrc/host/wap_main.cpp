// wap_tomo: the reconstruction driver, headless. It reads the reference driver's command
// line (Project/main.cpp:15-176), builds VolumeData and PSART as that driver does
// (Project/main.cpp:190-212) and runs PSART::Render on the GPU instead of handing the
// scene to a GLUT window. Exit status 0 on success, 1 on a command-line, file or device
// error.
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <sys/stat.h>

#include <chrono>
#include <vector>

#include "wap_classes.h"

int main(int argc, char** argv) {
  static char msg[1 << 16];
  WapCliConfig cfg;
  const int rc = wap_cli_parse(argc, (const char* const*)argv, &cfg, msg, sizeof(msg));
  if (rc == WAP_CLI_EXIT) {
    fputs(msg, stdout);
    return 0;
  }
  if (rc != WAP_OK) {
    fputs(msg, stdout);
    return 1;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || cfg.device < 0 || cfg.device >= ndev) {
    fprintf(stderr, "wap_tomo: no GPU %d\n", cfg.device);
    return 1;
  }
  if (cfg.total_proj < 1 || cfg.nframes < 1 || cfg.vol[0] < 1 || cfg.vol[1] < 1 ||
      cfg.vol[2] < 1 || cfg.img[0] < 1 || cfg.img[1] < 1) {
    fprintf(stderr, "wap_tomo: sizes, frames and projections must be positive\n");
    return 1;
  }
  (void)hipSetDevice(cfg.device);
  PSART::DefaultDevice = cfg.device;

  VolumeData* data = new VolumeData;
  data->setParas(cfg.vol, cfg.voxel_spacing, cfg.nframes, cfg.total_proj);
  data->ReadStructuredFile();

  std::vector<const char*> projs(WAP_CLI_MAXFILES, nullptr);
  for (int i = 0; i < cfg.n_proj_files; ++i) projs[i] = cfg.proj_files[i];
  PSART* scene = new PSART(cfg.sigma, cfg.tau, cfg.lambda, cfg.output, cfg.img, cfg.total_proj,
                           cfg.algo_iters, cfg.sart_iters, cfg.alpha, cfg.sid, projs.data(),
                           cfg.oxyz, cfg.sdd, cfg.det_spacing, cfg.prior, cfg.bp, cfg.vol,
                           cfg.nframes, cfg.start_deg, cfg.nrounds);
  int status = scene->LastError() ? 1 : 0;

  std::vector<double> angles(cfg.total_proj, 0.0);
  if (!status && cfg.use_lds) {
    if (wap_low_discrepancy_angles(cfg.total_proj, cfg.lds[0], cfg.lds[1], cfg.lds[2],
                                   angles.data()) != WAP_OK ||
        scene->SetAngles(angles.data(), cfg.total_proj) != WAP_OK)
      status = 1;
  } else if (!status) {
    // the constructor has read ours_foam_compress.txt from the working directory when it
    // exists; another list replaces it
    int nread = 0;
    struct stat st;
    if (stat(cfg.angles, &st) != 0) {
      fprintf(stderr, "wap_tomo: no angle list %s: every view at 0 degrees, as the reference\n",
              cfg.angles);
    } else if (wap_read_angles(cfg.angles, cfg.total_proj, angles.data(), &nread) != WAP_OK ||
               scene->SetAngles(angles.data(), cfg.total_proj) != WAP_OK) {
      fprintf(stderr, "wap_tomo: %s\n", wap_host_error());
      status = 1;
    } else if (nread < cfg.total_proj) {
      fprintf(stderr, "wap_tomo: %s holds %d of %d angles; the rest are 0\n", cfg.angles, nread,
              cfg.total_proj);
    }
  }
  if (!status) {
    scene->save_mode = cfg.save;
    scene->SetData(data);
    const auto t0 = std::chrono::steady_clock::now();
    scene->Render();
    const double s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    status = scene->LastError() ? 1 : 0;
    printf("wap_tomo: %s, %d x %d x %d volume, %d views, %d frames, Render %.3f s\n",
           status ? "failed" : "done", cfg.vol[0], cfg.vol[1], cfg.vol[2], cfg.total_proj,
           cfg.nframes, s);
  }
  delete scene;
  delete data;
  return status;
}
