// Headless counterpart of the reference application's main (cppvolrend/main.cpp:49-68):
// registers the techniques with the RenderingManager, loads a volume (.raw, .syn,
// .nrrd/.nhrd, .dat, or one entry of a dataset list) or makes a seeded synthetic one,
// loads a .tf1d transfer function, renders frames, writes the eval.csv line and a
// PFM/PPM of the last frame.
//
//   cvr_cli --technique rc1p|ebs|cone|nn|gt|dos
//           [--volume file | --datasets dir/list --dataset id|name | --synth n[,seed[,blobs]]]
//           [--tf f.tf1d] [--size WxH] [--frames N] [--csv eval.csv] [--pfm out.pfm]
//           [--ppm out.ppm] [--light x,y,z] [--multiscale single|multisample|down|up]
//           [--kernel box|hat|catmull-rom|mitchell-netravali|bspline3|omoms3]
//           [--camera-states f --camera-state id|name] [--light-sources f --light-source id|name]
//           [--check-inputs]
//
// --camera-states / --light-sources read the reference's preset files
// (data/#list_camera_states, data/#list_light_sources) and select one camera state and
// the first light of one light-source list (renderingmanager.cpp:132-151).
// --multiscale selects BaseVolumeRenderer's multi-scaling mode (volrenderbase.h:28-33):
// the technique renders at twice (multisample, down) or half (up) the --size and the
// frame is resampled to it with --kernel (default hat).
//
// --check-inputs parses the inputs and prints their summary without using the GPU.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/cvr_synth.h"
#include "../host/volrender.h"

static int usage() {
  std::fprintf(stderr,
               "usage: cvr_cli --technique rc1p|ebs|cone|nn|gt|dos [--volume file | --datasets dir/list "
               "--dataset id|name | --synth n[,seed[,blobs]]] "
               "[--tf f.tf1d] [--size WxH] [--frames N] [--csv f] [--pfm f] [--ppm f] [--light x,y,z] "
               "[--multiscale single|multisample|down|up] "
               "[--kernel box|hat|catmull-rom|mitchell-netravali|bspline3|omoms3] "
               "[--camera-states f --camera-state id|name] [--light-sources f --light-source id|name] "
               "[--check-inputs]\n");
  return 2;
}

int main(int argc, char** argv) {
  std::string tech = "rc1p", volume, tfpath, csv, pfm, ppm, synth = "64,1,8", datasets, dataset;
  std::string cam_file, cam_sel, light_file, light_sel;
  int W = 512, H = 512, frames = 1, ms_mode = 0, kernel = CVR_FILTER_HAT;
  const char* ms_names[] = {"single", "multisample", "down", "up"};
  const char* kernel_names[] = {"box", "hat", "catmull-rom", "mitchell-netravali", "bspline3", "omoms3"};
  auto pick = [](const std::string& v, const char* const* names, int n) {
    for (int k = 0; k < n; ++k)
      if (v == names[k]) return k;
    return -1;
  };
  bool check_only = false, have_light = false;
  std::array<float, 3> light{0, 0, 1};
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto next = [&]() -> std::string { return i + 1 < argc ? argv[++i] : ""; };
    if (a == "--technique") tech = next();
    else if (a == "--volume") volume = next();
    else if (a == "--datasets") datasets = next();
    else if (a == "--dataset") dataset = next();
    else if (a == "--synth") synth = next();
    else if (a == "--tf") tfpath = next();
    else if (a == "--size") { if (std::sscanf(next().c_str(), "%dx%d", &W, &H) != 2) return usage(); }
    else if (a == "--frames") frames = std::atoi(next().c_str());
    else if (a == "--csv") csv = next();
    else if (a == "--pfm") pfm = next();
    else if (a == "--ppm") ppm = next();
    else if (a == "--light") { have_light = std::sscanf(next().c_str(), "%f,%f,%f", &light[0], &light[1], &light[2]) == 3; }
    else if (a == "--multiscale") { if ((ms_mode = pick(next(), ms_names, 4)) < 0) return usage(); }
    else if (a == "--kernel") { if ((kernel = pick(next(), kernel_names, 6)) < 0) return usage(); }
    else if (a == "--camera-states") cam_file = next();
    else if (a == "--camera-state") cam_sel = next();
    else if (a == "--light-sources") light_file = next();
    else if (a == "--light-source") light_sel = next();
    else if (a == "--check-inputs") check_only = true;
    else return usage();
  }
  std::string err;
  // presets: an index or a name
  auto select = [](const std::string& sel, int n, int by_name) {
    if (by_name >= 0) return by_name;
    if (!sel.empty() && sel.find_first_not_of("0123456789") == std::string::npos) {
      const int i = std::atoi(sel.c_str());
      return i < n ? i : -1;
    }
    return sel.empty() ? 0 : -1;
  };
  const vis::CameraData* cam_state = nullptr;
  const vis::LightSourceData* light_src = nullptr;
  vis::CameraStateList cams;
  vis::LightSourceList lights;
  if (!cam_file.empty()) {
    if (!cams.ReadCameraStates(cam_file, &err)) { std::fprintf(stderr, "camera states: %s\n", err.c_str()); return 1; }
    const int i = select(cam_sel, cams.NumberOfCameraStates(), cams.FindCameraState(cam_sel));
    if (i < 0) { std::fprintf(stderr, "camera states: no state '%s'\n", cam_sel.c_str()); return 1; }
    cam_state = cams.GetCameraState(i);
    if (check_only) {
      const vis::CameraData& c = *cam_state;
      std::printf("camera_states=%d camera[%d]=%s type=%s eye=%g,%g,%g center=%g,%g,%g up=%g,%g,%g\n",
                  cams.NumberOfCameraStates(), i, c.cam_setup_name.c_str(), c.c_type == vis::CameraData::ARCBALL ? "ARCBALL" : "FLIGHT",
                  c.eye[0], c.eye[1], c.eye[2], c.center[0], c.center[1], c.center[2], c.up[0], c.up[1], c.up[2]);
    }
  }
  if (!light_file.empty()) {
    if (!lights.ReadLightSourceLists(light_file, &err)) { std::fprintf(stderr, "light sources: %s\n", err.c_str()); return 1; }
    const int i = select(light_sel, lights.NumberOfLists(), lights.FindList(light_sel));
    if (i < 0 || lights.GetList(i)->m_lightsources.empty()) {
      std::fprintf(stderr, "light sources: no list '%s' with a light\n", light_sel.c_str());
      return 1;
    }
    light_src = &lights.GetList(i)->m_lightsources[0];
    if (check_only) {
      const vis::LightSourceData& l = *light_src;
      const std::array<float, 3> f = l.Forward();
      std::printf("light_lists=%d light[%d]=%s lights=%d position=%g,%g,%g forward=%g,%g,%g up=%g,%g,%g right=%g,%g,%g "
                  "spot=%g\n",
                  lights.NumberOfLists(), i, lights.GetList(i)->l_name.c_str(), int(lights.GetList(i)->m_lightsources.size()),
                  l.position[0], l.position[1], l.position[2], f[0], f[1], f[2], l.y_axis[0], l.y_axis[1], l.y_axis[2],
                  l.x_axis[0], l.x_axis[1], l.x_axis[2], l.spot_light_angle);
    }
  }
  vis::StructuredGridVolume* vol = nullptr;
  if (!datasets.empty()) {
    vis::DataProvider dp;
    const size_t t = datasets.find_last_of('/');
    const std::string dir = t == std::string::npos ? "." : datasets.substr(0, t);
    const std::string list = t == std::string::npos ? datasets : datasets.substr(t + 1);
    if (!dp.SetStructuredGridFileList(dir, list, &err)) { std::fprintf(stderr, "datasets: %s\n", err.c_str()); return 1; }
    int id = dp.FindStructuredGridName(dataset);
    if (id < 0 && !dataset.empty() && dataset.find_first_not_of("0123456789") == std::string::npos) id = std::atoi(dataset.c_str());
    if (id < 0 || id >= dp.GetNumberOfStructuredGrids()) {
      std::fprintf(stderr, "datasets: no dataset '%s' among %d\n", dataset.c_str(), dp.GetNumberOfStructuredGrids());
      return 1;
    }
    if (check_only) {
      std::printf("datasets");
      for (const std::string& n : dp.GetStructuredGridNameList()) std::printf(" [%s]", n.c_str());
      std::printf("\n");
    }
    vol = dp.LoadStructuredGrid(unsigned(id), &err);
    if (!vol) { std::fprintf(stderr, "volume: %s\n", err.c_str()); return 1; }
  } else if (!volume.empty()) {
    vol = vis::VolumeReader::ReadStructuredVolume(volume, &err);
    if (!vol) { std::fprintf(stderr, "volume: %s\n", err.c_str()); return 1; }
  } else {
    int n = 64, seed = 1, blobs = 8;
    std::sscanf(synth.c_str(), "%d,%d,%d", &n, &seed, &blobs);
    std::vector<uint8_t> bytes(size_t(n) * n * n);
    if (cvr_synth_blobs(n, n, n, 0, blobs, 0.0, 4, uint64_t(seed), bytes.data()) != 0) return 1;
    vol = new vis::StructuredGridVolume("synthetic", n, n, n);
    vol->SetArrayData(std::move(bytes), 1);
  }
  vis::TransferFunction1D* tf = nullptr;
  if (!tfpath.empty()) {
    tf = vis::TransferFunctionReader::ReadTransferFunction(tfpath, &err);
    if (!tf) { std::fprintf(stderr, "transfer function: %s\n", err.c_str()); delete vol; return 1; }
  } else {
    tf = new vis::TransferFunction1D(255);
    tf->AddRGBControlPoint(1.0, 0.8, 0.6, 0);
    tf->AddRGBControlPoint(0.9, 0.3, 0.2, 255);
    tf->AddAlphaControlPoint(0.0, 0);
    tf->AddAlphaControlPoint(0.0, 40);
    tf->AddAlphaControlPoint(0.6, 255);
    tf->Build();
  }
  if (check_only) {
    // sum over voxels i (x fastest) of (i + 1) * value, modulo 2^64, identifies the data
    uint64_t wsum = 0;
    const size_t nv = size_t(vol->GetWidth()) * vol->GetHeight() * vol->GetDepth();
    const uint8_t* vb = static_cast<const uint8_t*>(vol->GetArrayData());
    for (size_t i = 0; i < nv; ++i) {
      const uint64_t v = vol->GetBytesPerValue() == 1 ? vb[i] : uint64_t(vb[2 * i]) | (uint64_t(vb[2 * i + 1]) << 8);
      wsum += uint64_t(i + 1) * v;
    }
    std::printf("volume %s %dx%dx%d bytes=%d diagonal=%.6f sample(0,0,0)=%.9f scale=%g,%g,%g wsum=%llu\n",
                vol->GetName().c_str(), vol->GetWidth(), vol->GetHeight(), vol->GetDepth(), vol->GetBytesPerValue(),
                vol->GetDiagonal(), vol->GetNormalizedSample(0, 0, 0), vol->GetScaleX(), vol->GetScaleY(),
                vol->GetScaleZ(), (unsigned long long)wsum);
    const std::vector<double>& t = tf->Table();
    std::printf("tf max_density=%d extinction_input=%d", tf->GetMaxDensity(), tf->IsExtinctionCoefficientInput() ? 1 : 0);
    for (int i : {0, tf->GetMaxDensity() / 2, tf->GetMaxDensity()})
      std::printf(" t[%d]=%.9f,%.9f,%.9f,%.9f", i, t[i * 4], t[i * 4 + 1], t[i * 4 + 2], t[i * 4 + 3]);
    std::printf(" extn(0.5)=%.9f\n", tf->GetExtN(0.5));
    delete vol;
    delete tf;
    return 0;
  }
  RenderingManager* rm = RenderingManager::Instance();
  rm->AddVolumeRenderer(new RayCasting1Pass());
  rm->AddVolumeRenderer(new RC1PExtinctionBasedShading());
  rm->AddVolumeRenderer(new RC1PConeTracedOcclusion());
  rm->AddVolumeRenderer(new RC1PLearnedConeOcclusion());
  rm->AddVolumeRenderer(new RC1PConeLightGroundTruthSteps());
  rm->AddVolumeRenderer(new RC1PConeTracingDirOcclusionShading());
  const char* names[] = {"rc1p", "ebs", "cone", "nn", "gt", "dos"};
  int id = -1;
  for (int i = 0; i < 6; ++i)
    if (tech == names[i]) id = i;
  if (id < 0) return usage();
  rm->InitData(vol, tf, W, H);
  if (cam_state) {
    rm->GetCamera().SetData(cam_state->eye, cam_state->center, cam_state->up);
    rm->UpdateLightSourceCameraVectors();
  }
  if (light_src) rm->GetRenderingParameters().SetLightSource(*light_src);
  if (have_light) rm->SetLightDirection(light);
  if (!rm->SetCurrentVolumeRenderer(id)) {
    std::fprintf(stderr, "init: %s\n", rm->GetVolumeRenderer(id)->GetLastError().c_str());
    RenderingManager::DestroyInstance();
    return 1;
  }
  if (id == 1) {  // EBS with a directional light, like the configuration it is benchmarked on
    cvr_ebs_params p;
    static_cast<GpuVolumeRenderer*>(rm->GetCurrentVolumeRenderer())->GetParams(&p, sizeof p);
    p.shadow_type = 1;
    static_cast<GpuVolumeRenderer*>(rm->GetCurrentVolumeRenderer())->SetParams(&p, sizeof p);
  }
  rm->GetCurrentVolumeRenderer()->SetImageKernelFilter(unsigned(kernel));
  rm->GetCurrentVolumeRenderer()->SetCurrentMultiScalingMode(ms_mode);
  double ms = 0;
  if (!rm->EvaluateFrames(frames, csv, &ms)) {
    std::fprintf(stderr, "render: %s\n", rm->GetCurrentVolumeRenderer()->GetLastError().c_str());
    RenderingManager::DestroyInstance();
    return 1;
  }
  std::printf("%s %s %dx%d (rendered at %dx%d, %s): %.3f ms/frame (host loop, frame copied back)\n",
              rm->GetCurrentVolumeRenderer()->GetAbbreviationName(), rm->GetCurrentVolumeRenderer()->GetName(), W, H,
              rm->GetCurrentVolumeRenderer()->GetRenderWidth(), rm->GetCurrentVolumeRenderer()->GetRenderHeight(),
              ms_names[ms_mode], ms);
  if (!pfm.empty() && !rm->SaveFramePFM(pfm)) std::fprintf(stderr, "could not write %s\n", pfm.c_str());
  if (!ppm.empty() && !rm->SaveFramePPM(ppm)) std::fprintf(stderr, "could not write %s\n", ppm.c_str());
  RenderingManager::DestroyInstance();
  return 0;
}
