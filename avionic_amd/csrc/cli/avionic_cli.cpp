// `avionic` command line: the reference CLI's reconstruction path
// (CUDA/src/main.cpp:332-947, options_parser.cpp:79-330) over the C-ABI of
// libavionic.
//
//   avionic [options] -d w:h:d:nRO:nEnc1:nEnc2:c:f <kdata.bin> <mask|traj.bin> <output.bin>
//
// Supported: methods TV, TVTEMP, TGV2, ICTV, ICTGV2 on Cartesian and (-n)
// non-Cartesian 2D+t data, and TGV2_3D on one Cartesian volume (depth > 1);
// coil maps (-s) and initial image (-u); the .cfg parameter file (-p) with
// the reference's keys; -i maxIt, -j stopPDGap, -a adaptlambda, -w density,
// -e extra results, -v verbose, -g debugstep, -o data normalisation.  Not in
// this build (it stops with a message): raw-data import (-r), coil-map
// estimation when -s is missing, and the other methods.
// --dry-run reads and checks every input and prints the resolved settings as
// JSON without touching the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/avionic.h"

namespace {

struct Dims {
  unsigned width = 0, height = 0, depth = 0, readouts = 0, enc1 = 0, enc2 = 0, coils = 0,
           frames = 0;
};

struct Options {
  std::string kdata, mask, output, params, sens, u0, dens, method = "ICTGV2";
  Dims dims;
  bool have_dims = false, nonuniform = false, extradata = false, adaptlambda = false,
       verbose = false, dry_run = false, normalize = false, rawdata = false, method_set = false;
  int maxIt = -1;  // -1: from the config file / defaults
  double stopPDGap = -1.0;
  int debugstep = 10;
  std::map<std::string, std::string> cfg;  // "section.key" -> value
};

[[noreturn]] void die(const std::string &msg, int code = 2) {
  std::fprintf(stderr, "avionic: %s\n", msg.c_str());
  std::exit(code);
}

void check(int rc, const char *what) {
  if (rc != AVIONIC_OK) die(std::string(what) + ": " + avionic_last_error(), 1);
}

std::string trim(const std::string &s) {
  const size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  const size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

std::string upper(std::string s) {
  for (char &c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

// options_parser.cpp:52-69: width:height:depth:readouts:encodings_y:encodings_z:coils:frames
Dims parse_dims(const std::string &s) {
  std::vector<unsigned> v;
  std::stringstream ss(s);
  std::string tok;
  while (std::getline(ss, tok, ':')) {
    char *end = nullptr;
    const unsigned long x = std::strtoul(tok.c_str(), &end, 10);
    if (tok.empty() || *end) die("invalid dimensions '" + s + "'");
    v.push_back((unsigned)x);
  }
  if (v.size() != 8) die("invalid dimensions '" + s + "' (want w:h:d:nRO:nEnc1:nEnc2:c:f)");
  Dims d;
  d.width = v[0]; d.height = v[1]; d.depth = v[2]; d.readouts = v[3];
  d.enc1 = v[4]; d.enc2 = v[5]; d.coils = v[6]; d.frames = v[7];
  return d;
}

// Keys a parameter file may set: the reference's configuration options and
// its command-line options by long name (options_parser.cpp:79-123,
// 173-321).  The reference reads the file with boost's parse_config_file,
// which throws on any other key (options_parser.cpp:393); so does this reader.
bool known_cfg_key(const std::string &key) {
  static const std::set<std::string> top = {
      "help", "debugstep", "verbose", "dims", "nonuniform", "normalize", "extradata",
      "adaptlambda", "dens", "params", "sens", "uzero", "labeldata", "rawdata", "forceOSRemoval",
      "tpat", "slice", "gpudevice", "dataBS", "maskBS", "finalOutBS", "method", "maxIt",
      "stopPDGap", "kdata", "mask", "output"};
  static const std::map<std::string, std::set<std::string>> sections = {
      {"coil", {"uH1mu", "uReg", "uNrIt", "uTau", "uSigma", "uSigmaTauRatio", "uAlpha0", "uAlpha1",
                "b1Reg", "b1NrIt", "b1Tau", "b1Sigma", "b1SigmaTauRatio", "b1FinalReg",
                "b1FinalNrIt"}},
      {"tv", {"dx", "dy", "dt", "sigma", "tau", "sigmaTauRatio", "timeSpaceWeight", "lambda"}},
      {"tvtemp", {"dt", "sigma", "tau", "sigmaTauRatio", "lambda"}},
      {"tgv2", {"dx", "dy", "dt", "sigma", "tau", "sigmaTauRatio", "timeSpaceWeight", "lambda",
                "alpha0", "alpha1"}},
      {"tgv2_3D", {"dx", "dy", "dz", "sigma", "tau", "sigmaTauRatio", "lambda", "alpha0", "alpha1"}},
      {"ictv", {"dx", "dy", "dt", "sigma", "tau", "sigmaTauRatio", "timeSpaceWeight", "lambda",
                "alpha1", "alpha", "timeSpaceWeight2", "dx2", "dy2", "dt2"}},
      {"ictgv2", {"dx", "dy", "dt", "sigma", "tau", "sigmaTauRatio", "timeSpaceWeight", "lambda",
                  "alpha0", "alpha1", "alpha", "timeSpaceWeight2", "dx2", "dy2", "dt2"}},
      {"asltgv2recon4D", {"dx", "dy", "dz", "dt", "sigma", "tau", "sigmaTauRatio",
                          "timeSpaceWeight", "lambda_l", "lambda_c", "alpha0", "alpha1", "alpha"}},
      {"gpunufft", {"kernelWidth", "sectorWidth", "osf"}},
      {"h1", {"mu", "relTol", "absTol", "maxIt"}},
      {"adaptlambda", {"k", "d"}}};
  const size_t dot = key.find('.');
  if (dot == std::string::npos) return top.count(key) != 0;
  auto it = sections.find(key.substr(0, dot));
  return it != sections.end() && it->second.count(key.substr(dot + 1)) != 0;
}

// Configuration file (boost program_options syntax, config/default.cfg):
// "key = value" lines, "[section]" headers, '#' comments.
void read_cfg(const std::string &path, std::map<std::string, std::string> &cfg) {
  std::ifstream in(path);
  if (!in) die("cannot open parameter file " + path);
  std::string line, section;
  while (std::getline(in, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    line = trim(line);
    if (line.empty()) continue;
    if (line.front() == '[') {
      if (line.back() != ']') die("malformed section header in " + path + ": " + line);
      section = trim(line.substr(1, line.size() - 2));
      continue;
    }
    const size_t eq = line.find('=');
    if (eq == std::string::npos) die("malformed line in " + path + ": " + line);
    const std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
    const std::string full = section.empty() ? key : section + "." + key;
    if (!known_cfg_key(full)) die("unrecognised option '" + full + "' in " + path);
    cfg[full] = val;
  }
}

Options parse_args(int argc, char **argv) {
  Options o;
  std::vector<std::string> pos;
  auto need = [&](int &i) -> std::string {
    if (i + 1 >= argc) die(std::string("option ") + argv[i] + " needs a value");
    return argv[++i];
  };
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "-h" || a == "--help") {
      std::printf(
          "Usage: avionic [options] -d w:h:d:nRO:nEnc1:nEnc2:c:f <kdata> <mask/traj> <output>\n"
          "  -m, --method TV|TVTEMP|TGV2|TGV2_3D|ICTV|ICTGV2  -i, --maxIt N  -j, --stopPDGap G\n"
          "  -p, --params file.cfg         -s, --sens b1.bin -u, --uzero u0.bin\n"
          "  -n, --nonuniform  -w, --dens dcf.bin  -a, --adaptlambda  -e, --extradata\n"
          "  -v, --verbose     -g, --debugstep N   -o, --normalize   --dry-run\n");
      std::exit(0);
    } else if (a == "-d" || a == "--dims") {
      o.dims = parse_dims(need(i));
      o.have_dims = true;
    } else if (a == "-m" || a == "--method") {
      o.method = upper(need(i));
      o.method_set = true;
    } else if (a == "-i" || a == "--maxIt") {
      o.maxIt = std::atoi(need(i).c_str());
    } else if (a == "-j" || a == "--stopPDGap") {
      o.stopPDGap = std::atof(need(i).c_str());
    } else if (a == "-p" || a == "--params") {
      o.params = need(i);
    } else if (a == "-s" || a == "--sens") {
      o.sens = need(i);
    } else if (a == "-u" || a == "--uzero") {
      o.u0 = need(i);
    } else if (a == "-w" || a == "--dens") {
      o.dens = need(i);
    } else if (a == "-g" || a == "--debugstep") {
      o.debugstep = std::atoi(need(i).c_str());
    } else if (a == "-n" || a == "--nonuniform") {
      o.nonuniform = true;
    } else if (a == "-e" || a == "--extradata") {
      o.extradata = true;
    } else if (a == "-a" || a == "--adaptlambda") {
      o.adaptlambda = true;
    } else if (a == "-v" || a == "--verbose") {
      o.verbose = true;
    } else if (a == "-o" || a == "--normalize") {
      o.normalize = true;
    } else if (a == "-r" || a == "--rawdata") {
      o.rawdata = true;
    } else if (a == "--dry-run") {
      o.dry_run = true;
    } else if (!a.empty() && a[0] == '-') {
      die("unknown or unsupported option " + a);
    } else {
      pos.push_back(a);
    }
  }
  if (o.rawdata) die("raw-data import (-r) is not part of this build", 3);
  if (pos.size() != 3) die("expected <kdata> <mask/traj> <output> (got " + std::to_string(pos.size()) + ")");
  o.kdata = pos[0];
  o.mask = pos[1];
  o.output = pos[2];
  if (!o.have_dims) die("-d w:h:d:nRO:nEnc1:nEnc2:c:f is required");
  if (!o.params.empty()) read_cfg(o.params, o.cfg);
  // the config file may also name the method / iterations (default.cfg:6-7)
  if (!o.method_set && o.cfg.count("method")) o.method = upper(o.cfg["method"]);
  return o;
}

avionic_method method_of(const std::string &m) {
  if (m == "TV") return AVIONIC_TV;
  if (m == "TGV2") return AVIONIC_TGV2;
  if (m == "ICTGV2") return AVIONIC_ICTGV2;
  if (m == "TVTEMP") return AVIONIC_TVTEMP;
  if (m == "ICTV") return AVIONIC_ICTV;
  if (m == "TGV2_3D") return AVIONIC_TGV2_3D;
  die("method " + m + " is not part of this build (TV, TVTEMP, TGV2, TGV2_3D, ICTV, ICTGV2)", 3);
}

// Config section keys of one method onto the parameter struct
// (options_parser.cpp:190-275).
void apply_cfg(const Options &o, const std::string &sec, avionic_params &p) {
  auto get = [&](const char *k, double &dst) {
    auto it = o.cfg.find(sec + "." + k);
    if (it != o.cfg.end()) dst = std::atof(it->second.c_str());
  };
  get("dx", p.dx); get("dy", p.dy); get("dt", p.dt); get("dz", p.dz);
  get("sigma", p.sigma); get("tau", p.tau); get("sigmaTauRatio", p.sigmaTauRatio);
  get("timeSpaceWeight", p.timeSpaceWeight); get("lambda", p.lambda);
  get("alpha0", p.alpha0); get("alpha1", p.alpha1); get("alpha", p.alpha);
  get("timeSpaceWeight2", p.timeSpaceWeight2);
  get("dx2", p.dx2); get("dy2", p.dy2); get("dt2", p.dt2);
  auto top = [&](const char *k, double &dst) {
    auto it = o.cfg.find(k);
    if (it != o.cfg.end()) dst = std::atof(it->second.c_str());
  };
  double maxIt = p.maxIt, stop = p.stopPDGap;
  top("maxIt", maxIt);
  top("stopPDGap", stop);
  p.maxIt = (unsigned)maxIt;
  p.stopPDGap = stop;
  if (o.maxIt >= 0) p.maxIt = (unsigned)o.maxIt;
  if (o.stopPDGap >= 0) p.stopPDGap = o.stopPDGap;
  // -a: lambda from the [adaptlambda] k, d of the file (default 0, 1:
  // options_parser.cpp:315-321); otherwise the configured lambda as is
  p.adaptLambda = o.adaptlambda ? 1 : 0;
  p.adaptLambdaK = 0.0;
  p.adaptLambdaD = 1.0;
  top("adaptlambda.k", p.adaptLambdaK);
  top("adaptlambda.d", p.adaptLambdaD);
}

// [coil] section of the parameter file (options_parser.cpp:173-189) over the
// InitParams defaults; coil.uSigmaTauRatio / b1SigmaTauRatio are accepted and,
// as in the reference, only InitParams applies them.
avionic_coil_params coil_params_of(const Options &o) {
  avionic_coil_params cp;
  check(avionic_default_coil_params(&cp), "coil parameters");
  auto get = [&](const char *k, double &dst) {
    auto it = o.cfg.find(std::string("coil.") + k);
    if (it != o.cfg.end()) dst = std::atof(it->second.c_str());
  };
  auto geti = [&](const char *k, unsigned &dst) {
    auto it = o.cfg.find(std::string("coil.") + k);
    if (it != o.cfg.end()) dst = (unsigned)std::atol(it->second.c_str());
  };
  get("uH1mu", cp.uH1mu); get("uReg", cp.uReg); geti("uNrIt", cp.uNrIt);
  get("uTau", cp.uTau); get("uSigma", cp.uSigma); get("uAlpha0", cp.uAlpha0);
  get("uAlpha1", cp.uAlpha1); get("b1Reg", cp.b1Reg); geti("b1NrIt", cp.b1NrIt);
  get("b1Tau", cp.b1Tau); get("b1Sigma", cp.b1Sigma); get("b1FinalReg", cp.b1FinalReg);
  geti("b1FinalNrIt", cp.b1FinalNrIt);
  return cp;
}

struct HostVec {
  std::vector<double> re, im;
  bool complex = false;
  size_t n = 0;
};

HostVec load(const std::string &path, const char *what) {
  HostVec v;
  int cplx = 0;
  check(avionic_bin_read(path.c_str(), nullptr, nullptr, 0, &v.n, &cplx), what);
  v.complex = cplx != 0;
  v.re.resize(v.n);
  v.im.resize(v.n);
  check(avionic_bin_read(path.c_str(), v.re.data(), v.im.data(), v.n, &v.n, &cplx), what);
  return v;
}

std::vector<float> to_c64(const HostVec &v) {
  std::vector<float> out(2 * v.n);
  for (size_t i = 0; i < v.n; ++i) {
    out[2 * i] = (float)v.re[i];
    out[2 * i + 1] = (float)v.im[i];
  }
  return out;
}

std::string dir_of(const std::string &path) {
  const size_t s = path.find_last_of('/');
  return s == std::string::npos ? std::string(".") : path.substr(0, s);
}

struct Dev {
  void *p = nullptr;
  explicit Dev(size_t bytes) {
    if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) die("device allocation failed", 1);
  }
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  float *f() const { return (float *)p; }
};

void upload(Dev &d, const std::vector<float> &h) {
  if (hipMemcpy(d.p, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice) != hipSuccess)
    die("host-to-device copy failed", 1);
}

// Magnitude of the root-sum-of-squares over coils of coil images produced
// one coil at a time by `image(c, out)` (out: HW complex values).
std::vector<float> rss_of(size_t C, size_t HW, const std::function<void(size_t, std::vector<float> &)> &image) {
  std::vector<float> u(HW, 0.f), img(2 * HW);
  for (size_t c = 0; c < C; ++c) {
    image(c, img);
    for (size_t i = 0; i < HW; ++i) {
      const float a = std::hypot(img[2 * i], img[2 * i + 1]);
      u[i] += a * a;
    }
  }
  for (float &v : u) v = std::sqrt(v);
  return u;
}

// Root-sum-of-squares of the time-averaged coil images of radial data
// (NoncartesianCoilConstruction::TimeAveragedReconstruction without phase,
// noncartesian_coil_construction.cpp:21-123) from the sqrt(density)-weighted
// data: every frame's points of a coil gridded together, kernel width 3,
// grid oversampling 2 (:84-86).
std::vector<float> radial_tar_rss(unsigned W, unsigned H, unsigned C, unsigned T, unsigned nfe,
                                  unsigned spf, const std::vector<float> &kdh,
                                  const std::vector<float> &maskh, const std::vector<float> &wh) {
  const size_t HW = (size_t)W * H;
  Dev dtraj(sizeof(float) * maskh.size()), dw(sizeof(float) * wh.size()), dk(sizeof(float) * kdh.size()),
      du(sizeof(float) * 2 * HW), dcrec(sizeof(float) * 2 * HW * C);
  upload(dtraj, maskh);
  upload(dw, wh);
  upload(dk, kdh);
  avionic_operator *op = nullptr;
  check(avionic_noncartesian_operator_create(W, H, C, T, nfe, spf, dtraj.f(), dw.f(), 3, 2.0, &op),
        "time-averaging operator");
  check(avionic_coil_time_averaged(op, dk.f(), du.f(), dcrec.f(), 0, nullptr), "time-averaged images");
  std::vector<float> u(2 * HW), rss(HW);
  if (hipMemcpy(u.data(), du.p, sizeof(float) * u.size(), hipMemcpyDeviceToHost) != hipSuccess)
    die("device-to-host copy failed", 1);
  avionic_operator_destroy(op);
  for (size_t i = 0; i < HW; ++i) rss[i] = u[2 * i];
  return rss;
}

// Normalisation factor of -o (main.cpp:575-579), computed from the prepared
// data.  Cartesian (RawDataPreparation::NormalizeCartData,
// raw_data_preparation.cpp:821-844): the time-averaged coil images
// F(sum_t d_t / max(1, sum_t M_t)) of a centred operator
// (cartesian_coil_construction.cpp:23-76; the operator of
// raw_data_preparation.cpp:371-375 is centred), their root-sum-of-squares, and
// 255 / median.  Radial (NormalizeNonCartData, :736-793): the adjoint
// gridding of every frame's points together into one image per coil
// (noncartesian_coil_construction.cpp:21-123; kernel width 3, grid
// oversampling 2, :84-86), its root-sum-of-squares, and 255 * frames / median.
float datanorm_of(bool nonuniform, unsigned W, unsigned H, unsigned C, unsigned T, size_t M,
                  unsigned nfe, unsigned spf, const std::vector<float> &kdh,
                  const std::vector<float> &maskh, const std::vector<float> &wh) {
  const size_t HW = (size_t)W * H;
  std::vector<float> ones(2 * HW, 0.f);
  for (size_t i = 0; i < HW; ++i) ones[2 * i] = 1.f;
  Dev dones(sizeof(float) * ones.size()), dimg(sizeof(float) * 2 * HW);
  upload(dones, ones);
  avionic_operator *op = nullptr;
  std::vector<float> u;
  auto download = [&](std::vector<float> &out) {
    if (hipMemcpy(out.data(), dimg.p, sizeof(float) * out.size(), hipMemcpyDeviceToHost) != hipSuccess)
      die("device-to-host copy failed", 1);
  };
  if (!nonuniform) {
    std::vector<float> cnt(HW, 0.f);
    for (size_t t = 0; t < T; ++t)
      for (size_t i = 0; i < HW; ++i) cnt[i] += maskh[t * HW + i];
    for (float &v : cnt) v = std::max(v, 1.f);
    check(avionic_cartesian_operator_create(W, H, 1, 1, nullptr, 1, &op), "normalisation operator");
    Dev davg(sizeof(float) * 2 * HW);
    std::vector<float> avg(2 * HW);
    u = rss_of(C, HW, [&](size_t c, std::vector<float> &img) {
      std::fill(avg.begin(), avg.end(), 0.f);
      for (size_t t = 0; t < T; ++t)
        for (size_t i = 0; i < 2 * HW; ++i) avg[i] += kdh[2 * (t * C + c) * HW + i];
      for (size_t i = 0; i < HW; ++i) {
        avg[2 * i] /= cnt[i];
        avg[2 * i + 1] /= cnt[i];
      }
      upload(davg, avg);
      check(avionic_operator_forward(op, davg.f(), dimg.f(), dones.f(), nullptr), "time-averaged image");
      download(img);
    });
  } else {
    u = radial_tar_rss(W, H, C, T, nfe, spf, kdh, maskh, wh);
  }
  if (op) avionic_operator_destroy(op);
  float median = 0.f;
  check(avionic_normalization_factor(u.data(), u.size(), &median), "normalisation factor");
  return nonuniform ? 255.f * (float)T / median : 255.f / median;
}

void export_cb(const char *name, const float *data, size_t count, void *user) {
  const std::string path = *(const std::string *)user + "/" + name;
  std::vector<double> re(count), im(count);
  for (size_t i = 0; i < count; ++i) {
    re[i] = data[2 * i];
    im[i] = data[2 * i + 1];
  }
  if (avionic_bin_write(path.c_str(), re.data(), im.data(), count) != AVIONIC_OK)
    std::fprintf(stderr, "avionic: could not write %s: %s\n", path.c_str(), avionic_last_error());
}

}  // namespace

int main(int argc, char **argv) {
  Options o = parse_args(argc, argv);
  const Dims &d = o.dims;
  const avionic_method method = method_of(o.method);
  // TGV2_3D reconstructs one volume: the third axis is the depth, k-space and
  // coil maps are coil-major [c][depth][h][w] (main.cpp:403-406, 466-474)
  const bool vol = method == AVIONIC_TGV2_3D;
  if (!vol && d.depth > 1) die("3D data (depth > 1) needs -m TGV2_3D", 3);
  if (vol && o.nonuniform) die("non-Cartesian 3D data is not part of this build", 3);
  if (vol && o.normalize) die("data normalisation (-o) of 3D data is not part of this build", 3);
  if (vol && d.frames != 1) die("TGV2_3D takes one frame");
  const unsigned W = d.width, H = d.height, C = d.coils, T = vol ? d.depth : d.frames;
  if (!W || !H || !C || !T) die("width, height, coils and frames must be positive");
  const size_t HW = (size_t)W * H, N = HW * T;
  const size_t M = o.nonuniform ? (size_t)d.readouts * d.enc1 : 0;
  const size_t K = o.nonuniform ? M * C * T : HW * C * T;

  avionic_params p;
  check(avionic_default_params(method, &p), "default parameters");
  const char *sec = method == AVIONIC_TV       ? "tv"
                    : method == AVIONIC_TVTEMP ? "tvtemp"
                    : method == AVIONIC_TGV2   ? "tgv2"
                    : method == AVIONIC_ICTV   ? "ictv"
                    : method == AVIONIC_TGV2_3D ? "tgv2_3D"
                                               : "ictgv2";
  apply_cfg(o, sec, p);
  if (!p.adaptLambda && !(p.lambda > 0.0))
    die("no lambda: set " + std::string(sec) + ".lambda in the parameter file (-p) or pass -a");
  int kw = 3;
  double osf = 2.0;  // options_parser.cpp:300-309 defaults
  if (o.cfg.count("gpunufft.kernelWidth")) kw = (int)std::atof(o.cfg["gpunufft.kernelWidth"].c_str());
  if (o.cfg.count("gpunufft.osf")) osf = std::atof(o.cfg["gpunufft.osf"].c_str());

  if (vol && o.sens.empty())  // main.cpp:621-626
    die("coil construction for 3D reconstruction is not implemented: provide the coil maps (-s)", 3);

  // ---- inputs
  HostVec kd = load(o.kdata, "k-space data");
  if (kd.n != K) die("k-space data has " + std::to_string(kd.n) + " values, dims imply " + std::to_string(K));
  HostVec mk = load(o.mask, o.nonuniform ? "trajectory" : "mask");
  const size_t mask_n = o.nonuniform ? 2 * M * T : N;
  if (mk.n != mask_n)
    die(std::string(o.nonuniform ? "trajectory" : "mask") + " has " + std::to_string(mk.n) +
        " values, dims imply " + std::to_string(mask_n));
  HostVec b1, u0, w;
  const bool have_b1 = !o.sens.empty();
  const size_t B1N = vol ? N * C : HW * C;
  if (have_b1) {
    b1 = load(o.sens, "coil maps");
    if (b1.n != B1N)
      die("coil maps have " + std::to_string(b1.n) + " values, dims imply " + std::to_string(B1N));
  }
  const avionic_coil_params coilp = coil_params_of(o);
  if (!o.u0.empty()) {
    u0 = load(o.u0, "initial image");
    if (vol ? u0.n != N : (u0.n != HW && u0.n != N))
      die(vol ? "initial image must hold w*h*depth values"
              : "initial image must hold w*h or w*h*frames values");
  }
  if (o.nonuniform) {
    if (o.dens.empty()) die("non-Cartesian data needs the density compensation (-w)");
    w = load(o.dens, "density compensation");
    if (w.n != M * T) die("density compensation has " + std::to_string(w.n) + " values, dims imply " + std::to_string(M * T));
  }

  if (o.dry_run) {
    std::printf(
        "{\"method\": \"%s\", \"width\": %u, \"height\": %u, \"coils\": %u, \"frames\": %u, "
        "\"nonuniform\": %s, \"points_per_frame\": %zu, \"maxIt\": %u, \"stopPDGap\": %g, "
        "\"sigma\": %g, \"tau\": %g, \"sigmaTauRatio\": %g, \"timeSpaceWeight\": %g, "
        "\"timeSpaceWeight2\": %g, \"lambda\": %g, \"adaptLambda\": %d, \"adaptLambdaK\": %g, "
        "\"adaptLambdaD\": %g, \"alpha0\": %g, \"alpha1\": %g, \"alpha\": %g, "
        "\"kernelWidth\": %d, \"osf\": %g, \"initial_image\": \"%s\", \"coil_maps\": \"%s\", "
        "\"normalize\": %s, \"output\": \"%s\"}\n",
        o.method.c_str(), W, H, C, T, o.nonuniform ? "true" : "false", M, p.maxIt, p.stopPDGap,
        p.sigma, p.tau, p.sigmaTauRatio, p.timeSpaceWeight, p.timeSpaceWeight2, p.lambda,
        p.adaptLambda, p.adaptLambdaK, p.adaptLambdaD, p.alpha0, p.alpha1, p.alpha, kw, osf,
        !o.u0.empty() ? "file" : (vol ? "zero" : (have_b1 ? "time-averaged" : "coil-construction")),
        have_b1 ? "file" : "coil-construction", o.normalize ? "true" : "false",
        o.output.c_str());
    return 0;
  }

  // ---- data preparation (main.cpp:509-579)
  std::vector<float> kdh = to_c64(kd);
  std::vector<float> maskh(mk.n);
  for (size_t i = 0; i < mk.n; ++i) maskh[i] = (float)mk.re[i];
  std::vector<float> wh;
  if (vol) {
    // zero the unsampled k-space of every coil volume (main.cpp:466-474)
    for (size_t c = 0; c < C; ++c)
      for (size_t i = 0; i < N; ++i) {
        kdh[2 * (c * N + i)] *= maskh[i];
        kdh[2 * (c * N + i) + 1] *= maskh[i];
      }
  } else if (!o.nonuniform) {
    // zero the unsampled k-space of every frame and coil (main.cpp:509-519)
    for (size_t t = 0; t < T; ++t)
      for (size_t c = 0; c < C; ++c)
        for (size_t i = 0; i < HW; ++i) {
          const float m = maskh[t * HW + i];
          float *z = &kdh[2 * ((t * C + c) * HW + i)];
          z[0] *= m;
          z[1] *= m;
        }
  } else {
    // sqrt(density) weighting of the data (main.cpp:546-570); the operator
    // applies sqrt(density) to its own output
    wh.resize(w.n);
    for (size_t i = 0; i < w.n; ++i) wh[i] = (float)w.re[i];
    for (size_t t = 0; t < T; ++t)
      for (size_t c = 0; c < C; ++c)
        for (size_t m = 0; m < M; ++m) {
          const float s = std::sqrt(std::max(0.f, wh[t * M + m]));
          float *z = &kdh[2 * ((t * C + c) * M + m)];
          z[0] *= s;
          z[1] *= s;
        }
  }

  // ---- normalisation (-o, main.cpp:575-579): scale the prepared data
  float datanorm = 1.f;
  if (o.normalize) {
    datanorm = datanorm_of(o.nonuniform, W, H, C, T, M, d.readouts, d.enc1, kdh, maskh, wh);
    for (float &v : kdh) v *= datanorm;
    if (o.verbose) std::printf("avionic: datanorm factor %g\n", datanorm);
  }

  Dev dk(sizeof(float) * kdh.size()), dmask(sizeof(float) * maskh.size()),
      db1(sizeof(float) * 2 * B1N), dx(sizeof(float) * 2 * N), dw(sizeof(float) * (wh.size() + 1));
  upload(dk, kdh);
  upload(dmask, maskh);
  if (have_b1) upload(db1, to_c64(b1));
  if (!wh.empty()) upload(dw, wh);

  avionic_operator *op = nullptr;
  if (o.nonuniform)
    check(avionic_noncartesian_operator_create(W, H, C, T, d.readouts, d.enc1, dmask.f(), dw.f(), kw,
                                               osf, &op),
          "non-Cartesian operator");
  else if (vol)
    check(avionic_cartesian3d_operator_create(W, H, T, C, dmask.f(), &op), "3D Cartesian operator");
  else
    check(avionic_cartesian_operator_create(W, H, C, T, dmask.f(), 0, &op), "Cartesian operator");

  // ---- coil maps and initial image (main.cpp:592-654).  Both derive from
  // the time-averaged coil images crec_c: Cartesian, the uncentred masked
  // operator's TimeAveragedReconstruction (cartesian_coil_construction.cpp
  // :23-87, the operator of main.cpp:199-200); radial, every frame's points
  // gridded together (noncartesian_coil_construction.cpp:21-123).  The
  // library's avionic_coil_time_averaged computes both.
  //   no -s: coil construction (coil_construction.cpp:447-557); u0 = the
  //     image estimate of the last coil step (main.cpp:639), divided by the
  //     frame count for radial data (:641-644); u0_reconstructed.bin and
  //     b1_reconstructed.bin are written next to the output (:645-648);
  //   -s without -u: u0 = (|crec| + sum_c conj(b1_c) crec_c) / 2, the
  //     root-sum-of-squares without phase (PerformInitalizationGivenB1,
  //     main.cpp:250-287; the reference builds its radial operator there for
  //     Cartesian data too, here Cartesian data keeps its own operator);
  //   -s and -u: u0 from the file, one frame replicated (main.cpp:772-778).
  std::vector<float> x0(2 * N, 0.f), u1(2 * HW, 0.f);
  const bool need_tar = !vol && (!have_b1 || o.u0.empty());
  std::vector<float> crec_h, tu_h(2 * HW);
  Dev dcrec(sizeof(float) * 2 * HW * C), dtu(sizeof(float) * 2 * HW);
  if (need_tar) {
    // PerformCoilConstruction takes the images with phase, GivenB1 without
    check(avionic_coil_time_averaged(op, dk.f(), dtu.f(), dcrec.f(), have_b1 ? 0 : 1, nullptr),
          "time-averaged reconstruction");
    crec_h.resize(2 * HW * C);
    if (hipMemcpy(crec_h.data(), dcrec.p, sizeof(float) * crec_h.size(), hipMemcpyDeviceToHost) !=
            hipSuccess ||
        hipMemcpy(tu_h.data(), dtu.p, sizeof(float) * tu_h.size(), hipMemcpyDeviceToHost) != hipSuccess)
      die("device-to-host copy failed", 1);
  }
  if (!have_b1) {
    Dev du(sizeof(float) * 2 * HW * C);
    check(avionic_coil_construction_from_images(W, H, C, dcrec.f(), dtu.f(), &coilp, du.f(), db1.f(),
                                                nullptr, nullptr),
          "coil construction");
    std::vector<float> uall(2 * HW * C), b1h(2 * HW * C);
    if (hipMemcpy(uall.data(), du.p, sizeof(float) * uall.size(), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(b1h.data(), db1.p, sizeof(float) * b1h.size(), hipMemcpyDeviceToHost) != hipSuccess)
      die("device-to-host copy failed", 1);
    const float s = o.nonuniform ? 1.f / (float)T : 1.f;
    for (size_t i = 0; i < 2 * HW; ++i) u1[i] = uall[2 * (C - 1) * HW + i] * s;
    const std::string dir = dir_of(o.output);
    auto write_c = [&](const char *name, const std::vector<float> &v) {
      const size_t n = v.size() / 2;
      std::vector<double> re(n), im(n);
      for (size_t i = 0; i < n; ++i) {
        re[i] = v[2 * i];
        im[i] = v[2 * i + 1];
      }
      check(avionic_bin_write((dir + "/" + name).c_str(), re.data(), im.data(), n), name);
    };
    write_c("u0_reconstructed.bin", u1);
    write_c("b1_reconstructed.bin", b1h);
    if (o.verbose) std::printf("avionic: coil maps estimated (%u coils)\n", C);
  } else if (!vol && o.u0.empty()) {
    const std::vector<float> bh = to_c64(b1);
    for (size_t i = 0; i < HW; ++i) {
      double re = 0.0, im = 0.0;
      for (size_t c = 0; c < C; ++c) {
        const double br = bh[2 * (c * HW + i)], bi = bh[2 * (c * HW + i) + 1];
        const double cr = crec_h[2 * (c * HW + i)], ci = crec_h[2 * (c * HW + i) + 1];
        re += br * cr + bi * ci;  // conj(b1) crec
        im += br * ci - bi * cr;
      }
      u1[2 * i] = (float)(0.5 * (tu_h[2 * i] + re));
      u1[2 * i + 1] = (float)(0.5 * (tu_h[2 * i + 1] + im));
    }
  }
  if (vol) {
    // the volume from the file (main.cpp:770-773), or zero
    if (!o.u0.empty()) x0 = to_c64(u0);
  } else if (have_b1 && !o.u0.empty()) {
    const std::vector<float> u = to_c64(u0);
    for (size_t t = 0; t < T; ++t)
      for (size_t i = 0; i < 2 * HW; ++i) x0[2 * t * HW + i] = u[(u0.n == N ? 2 * t * HW : 0) + i];
  } else {
    for (size_t t = 0; t < T; ++t) std::memcpy(&x0[2 * t * HW], u1.data(), sizeof(float) * 2 * HW);
  }
  upload(dx, x0);

  // ---- reconstruction (main.cpp:780-789)
  avionic_recon *rec = nullptr;
  check(avionic_recon_create(method, W, H, C, T, &p, op, &rec), "solver");
  check(avionic_recon_set_verbose(rec, o.verbose ? 1 : 0), "verbose");
  check(avionic_recon_set_debug(rec, o.debugstep > 0 ? 1 : 0, o.debugstep), "debug");
  if (o.verbose) std::printf("avionic: %s %ux%u x %u frames x %u coils, %u iterations\n",
                             o.method.c_str(), W, H, T, C, p.maxIt);
  check(avionic_recon_run(rec, dk.f(), dx.f(), db1.f(), nullptr), "reconstruction");

  // ---- output (main.cpp:831-839, 919-921)
  std::vector<float> xh(2 * N);
  if (hipMemcpy(xh.data(), dx.p, sizeof(float) * xh.size(), hipMemcpyDeviceToHost) != hipSuccess)
    die("device-to-host copy failed", 1);
  std::vector<double> re(N), im(N);
  for (size_t i = 0; i < N; ++i) {
    re[i] = xh[2 * i];
    im[i] = xh[2 * i + 1];
  }
  check(avionic_bin_write(o.output.c_str(), re.data(), im.data(), N), "writing the result");
  // the factor next to the output, as a one-element complex vector
  // (main.cpp:792-796)
  if (o.normalize) {
    const double re1 = datanorm, im1 = 0.0;
    check(avionic_bin_write((dir_of(o.output) + "/datanorm_factor.bin").c_str(), &re1, &im1, 1),
          "writing datanorm_factor.bin");
  }
  if (o.extradata) {
    std::string dir = dir_of(o.output);
    check(avionic_recon_export(rec, export_cb, &dir), "exporting additional results");
  }
  avionic_recon_destroy(rec);
  avionic_operator_destroy(op);
  return 0;
}
