// Headless counterpart of the reference application's main (cppvolrend/main.cpp:49-68):
// registers the techniques with the RenderingManager, loads a volume (.raw, .syn,
// .nrrd/.nhrd, .dat, or one entry of a dataset list) or makes a seeded synthetic one,
// loads a .tf1d transfer function, renders frames, writes the eval.csv line and a
// PFM/PPM of the last frame.
//
//   cvr_cli --technique rc1p|ebs|cone|nn|gt|dos
//           [--volume file | --datasets dir/list --dataset id|name | --synth n[,seed[,blobs]]]
//           [--tf f.tf1d] [--size WxH] [--frames N] [--csv eval.csv] [--pfm out.pfm]
//           [--ppm out.ppm] [--light x,y,z] [--check-inputs]
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
               "[--check-inputs]\n");
  return 2;
}

int main(int argc, char** argv) {
  std::string tech = "rc1p", volume, tfpath, csv, pfm, ppm, synth = "64,1,8", datasets, dataset;
  int W = 512, H = 512, frames = 1;
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
    else if (a == "--check-inputs") check_only = true;
    else return usage();
  }
  std::string err;
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
  double ms = 0;
  if (!rm->EvaluateFrames(frames, csv, &ms)) {
    std::fprintf(stderr, "render: %s\n", rm->GetCurrentVolumeRenderer()->GetLastError().c_str());
    RenderingManager::DestroyInstance();
    return 1;
  }
  std::printf("%s %s %dx%d: %.3f ms/frame (host loop, frame copied back)\n",
              rm->GetCurrentVolumeRenderer()->GetAbbreviationName(), rm->GetCurrentVolumeRenderer()->GetName(), W, H,
              ms);
  if (!pfm.empty() && !rm->SaveFramePFM(pfm)) std::fprintf(stderr, "could not write %s\n", pfm.c_str());
  if (!ppm.empty() && !rm->SaveFramePPM(ppm)) std::fprintf(stderr, "could not write %s\n", ppm.c_str());
  RenderingManager::DestroyInstance();
  return 0;
}
