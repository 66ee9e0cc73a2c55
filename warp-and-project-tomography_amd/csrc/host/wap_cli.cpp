// Command line of the reconstruction driver (wap_tomo), with the flags, defaults and
// checks of the reference's argtable2 table (Project/main.cpp:15-176):
//   - short options take their value attached (-s0.5) or as the next word (-s 0.5); long
//     options as --Sigma=0.5 or --Sigma 0.5, and a unique prefix of a long name works;
//   - options with several values (-f, -u, -i, --oxyz, --sdg) are repeated, one value
//     each; more than the table's maximum is an error, and so is none where the table's
//     minimum is 1;
//   - integers are decimal, 0x hex, 0o octal or 0b binary with an optional KB/MB/GB
//     suffix; doubles are strtod numbers;
//   - --help wins over --version, which wins over errors; all errors are reported.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include <string>
#include <vector>

#include "../../../include/wap.h"

namespace {

enum Kind { K_DBL, K_INT, K_FILE, K_STR, K_LIT };

struct Opt {
  char shortc;       // 0 = none
  const char* name;  // long name (nullptr = none)
  Kind kind;
  int mincount, maxcount;
  const char* datatype;
  const char* glossary;
  int count = 0;
};

// The table, in the reference's order; the last entries have no reference counterpart.
enum {
  O_SIGMA, O_TAU, O_LAMBDA, O_ALPHA, O_OUT, O_PROJS, O_VOL, O_IMG, O_VS, O_DS, O_SID, O_SDD,
  O_OXYZ, O_SDG, O_NFRAMES, O_NROUNDS, O_ALGO, O_SART, O_PRIOR, O_BP, O_HELP, O_VERSION,
  O_NPROJ, O_ANGLES, O_LDS, O_SAVE, O_DEVICE, O_COUNT
};

std::vector<Opt> make_table() {
  return {
      {'s', "Sigma", K_DBL, 0, 1, "<double>", "TV weight (default 0.3)"},
      {'t', "Tau", K_DBL, 0, 1, "<double>", "temporal-prior weight (default 0.3)"},
      {'l', "Lambda", K_DBL, 0, 1, "<double>", "trade-off weight (default 0.1; not used)"},
      {'a', "Alpha", K_DBL, 0, 1, "<double>", "SART relaxation (default 0.1)"},
      {'o', nullptr, K_FILE, 0, 1, "<output>", "output prefix (default Test/test)"},
      {'i', "projsvolume", K_FILE, 1, 20, "<file>",
       "projection stack (.hdr/.img/.nii, .tif, .mha/.mhd, raw float32)"},
      {'f', "XYZ", K_INT, 1, 3, "<int>", "volume size W H D, one per flag (rest 63)"},
      {'u', "imgUV", K_INT, 1, 2, "<int>", "detector W H, one per flag (rest 512)"},
      {'c', "voxelspacing", K_DBL, 0, 1, "<double>", "voxel spacing (default 1)"},
      {'d', "detector spacing", K_DBL, 0, 1, "<double>", "detector spacing (default 1)"},
      {'g', "sid", K_DBL, 0, 1, "<double>", "source to rotation centre (default 400)"},
      {'j', "sdd", K_DBL, 0, 1, "<double>", "source to detector (default 200)"},
      {0, "oxyz", K_DBL, 1, 3, "<double>", "volume centre offset x y z, one per flag"},
      {0, "sdg", K_DBL, 1, 20, "<double>", "start degree of each sequence, one per flag"},
      {'v', "nframes", K_INT, 0, 1, "<int>", "number of frames (default 1)"},
      {'r', "nframes", K_INT, 0, 1, "<int>", "rounds per frame (default 1)"},
      {'b', "AlgorithmIter", K_INT, 0, 1, "<int>", "algorithm iterations (default 2; not used)"},
      {'p', "SartIter", K_INT, 0, 1, "<int>", "SART nested iterations (default 2; not used)"},
      {0, "prior", K_STR, 0, 1, "{STV,SAD,ATV}", "prior (default ATV; not used)"},
      {0, "bp", K_STR, 0, 1, "{Voxelbased,Raybased}", "backprojector (default Voxelbased)"},
      {'h', "help", K_LIT, 0, 1, "", "print this help and exit"},
      {0, "version", K_LIT, 0, 1, "", "print version information and exit"},
      {0, "nproj", K_INT, 0, 1, "<int>", "projections in the stack (default 192)"},
      {0, "angles", K_FILE, 0, 1, "<file>", "angle list (default ours_foam_compress.txt)"},
      {0, "lds", K_STR, 0, 1, "<Ns,p,Nz>", "low-discrepancy angles instead of the list"},
      {0, "save", K_STR, 0, 1, "{none,final,all}", "volumes written (default final)"},
      {0, "device", K_INT, 0, 1, "<int>", "GPU ordinal (default 0)"},
  };
}

// argtable2's integer syntax: optional sign, 0x/0o/0b prefixes, KB/MB/GB suffixes.
bool parse_int(const char* s, int* out) {
  const char* p = s;
  while (*p == ' ' || *p == '\t') ++p;
  bool neg = false;
  if (*p == '+' || *p == '-') neg = *p++ == '-';
  int base = 10;
  if (p[0] == '0' && (p[1] == 'x' || p[1] == 'X')) base = 16, p += 2;
  else if (p[0] == '0' && (p[1] == 'o' || p[1] == 'O')) base = 8, p += 2;
  else if (p[0] == '0' && (p[1] == 'b' || p[1] == 'B')) base = 2, p += 2;
  if (*p == 0 || *p == '-' || *p == '+') return false;
  char* end;
  long long v = strtoll(p, &end, base);
  if (end == p) return false;
  if (strcasecmp(end, "KB") == 0) v *= 1024LL, end += 2;
  else if (strcasecmp(end, "MB") == 0) v *= 1024LL * 1024, end += 2;
  else if (strcasecmp(end, "GB") == 0) v *= 1024LL * 1024 * 1024, end += 2;
  if (*end) return false;
  if (neg) v = -v;
  if (v < -2147483648LL || v > 2147483647LL) return false;
  *out = (int)v;
  return true;
}

bool parse_dbl(const char* s, double* out) {
  char* end;
  const double v = strtod(s, &end);
  if (end == s || *end) return false;
  *out = v;
  return true;
}

void copy_path(char* dst, const char* src) {
  strncpy(dst, src, WAP_CLI_PATH - 1);
  dst[WAP_CLI_PATH - 1] = 0;
}

std::string label(const Opt& o) {
  std::string s;
  if (o.shortc) s += std::string("-") + o.shortc;
  if (o.name) s += std::string(s.empty() ? "" : "|") + "--" + o.name;
  return s;
}

void defaults(WapCliConfig* c) {
  memset(c, 0, sizeof(*c));
  copy_path(c->output, "Test/test");
  c->sigma = 0.3;
  c->tau = 0.3;
  c->lambda = 0.1;
  c->algo_iters = 2;
  c->sart_iters = 2;
  c->alpha = 0.1;
  c->voxel_spacing = 1.0;
  c->sdd = 200.0;
  c->vol[0] = c->vol[1] = c->vol[2] = 63;
  c->img[0] = c->img[1] = 512;
  strcpy(c->prior, "ATV");
  strcpy(c->bp, "Voxelbased");
  c->sid = 400.0;
  copy_path(c->proj_files[0], "proj");
  c->det_spacing = 1.0;
  c->nframes = 1;
  c->nrounds = 1;
  c->total_proj = 192;
  copy_path(c->angles, "ours_foam_compress.txt");
  c->save = 1;
}

// Stores one value of option k; returns false with a reason when the value is invalid.
bool store(WapCliConfig* c, int k, int idx, const char* v, std::string& why) {
  int iv = 0;
  double dv = 0;
  switch (k) {
    case O_SIGMA: case O_TAU: case O_LAMBDA: case O_ALPHA: case O_VS: case O_DS: case O_SID:
    case O_SDD: case O_OXYZ: case O_SDG:
      if (!parse_dbl(v, &dv)) return false;
      if (k == O_SIGMA) c->sigma = dv;
      if (k == O_TAU) c->tau = dv;
      if (k == O_LAMBDA) c->lambda = dv;
      if (k == O_ALPHA) c->alpha = dv;
      if (k == O_VS) c->voxel_spacing = dv;
      if (k == O_DS) c->det_spacing = dv;
      if (k == O_SID) c->sid = dv;
      if (k == O_SDD) c->sdd = dv;
      if (k == O_OXYZ) c->oxyz[idx] = dv;
      if (k == O_SDG) c->start_deg[idx] = dv, c->n_start_deg = idx + 1;
      return true;
    case O_VOL: case O_IMG: case O_NFRAMES: case O_NROUNDS: case O_ALGO: case O_SART:
    case O_NPROJ: case O_DEVICE:
      if (!parse_int(v, &iv)) return false;
      if (k == O_VOL) c->vol[idx] = iv;
      if (k == O_IMG) c->img[idx] = iv;
      if (k == O_NFRAMES) c->nframes = iv;
      if (k == O_NROUNDS) c->nrounds = iv;
      if (k == O_ALGO) c->algo_iters = iv;
      if (k == O_SART) c->sart_iters = iv;
      if (k == O_NPROJ) c->total_proj = iv;
      if (k == O_DEVICE) c->device = iv;
      return true;
    case O_OUT:
      copy_path(c->output, v);
      return true;
    case O_PROJS:
      copy_path(c->proj_files[idx], v);
      c->n_proj_files = idx + 1;
      return true;
    case O_ANGLES:
      copy_path(c->angles, v);
      return true;
    case O_PRIOR: case O_BP:
      if (strlen(v) >= sizeof(c->prior)) {
        why = "value too long";
        return false;
      }
      strcpy(k == O_PRIOR ? c->prior : c->bp, v);
      return true;
    case O_LDS: {
      int a, b, d;
      char tail;
      if (sscanf(v, "%d,%d,%d%c", &a, &b, &d, &tail) != 3 || a < 1 || b < 2 || d < 0) {
        why = "expected Ns,p,Nz with Ns >= 1, p >= 2, Nz >= 0";
        return false;
      }
      c->lds[0] = a, c->lds[1] = b, c->lds[2] = d, c->use_lds = 1;
      return true;
    }
    case O_SAVE:
      if (!strcmp(v, "none")) c->save = 0;
      else if (!strcmp(v, "final")) c->save = 1;
      else if (!strcmp(v, "all")) c->save = 2;
      else {
        why = "expected none, final or all";
        return false;
      }
      return true;
    default:
      return false;
  }
}

std::string help_text(const char* prog, const std::vector<Opt>& t) {
  std::string s = std::string("Usage: ") + prog;
  for (const Opt& o : t) {
    s += " [" + label(o);
    if (o.kind != K_LIT) s += (o.shortc && !o.name ? " " : "=") + std::string(o.datatype);
    s += "]";
    if (o.maxcount > 1) s += "...";
  }
  s += "\nMI355X warp-and-project reconstruction (headless PSART::Render).\n";
  for (const Opt& o : t) {
    std::string l = label(o);
    if (o.kind != K_LIT) l += (o.shortc && !o.name ? " " : "=") + std::string(o.datatype);
    char line[512];
    snprintf(line, sizeof(line), "  %-34s %s%s\n", l.c_str(), o.glossary,
             o.mincount > 0 ? " [required]" : "");
    s += line;
  }
  return s;
}

void put(char* msg, size_t len, const std::string& s) {
  if (!msg || len == 0) return;
  const size_t n = std::min(len - 1, s.size());
  memcpy(msg, s.data(), n);
  msg[n] = 0;
}

}  // namespace

extern "C" int wap_cli_parse(int argc, const char* const* argv, WapCliConfig* cfg, char* msg,
                             size_t msglen) {
  if (!cfg || argc < 0 || (argc > 0 && !argv)) return WAP_E_ARG;
  put(msg, msglen, "");
  defaults(cfg);
  std::vector<Opt> t = make_table();
  const char* prog = argc > 0 && argv[0] ? argv[0] : "wap_tomo";
  std::vector<std::string> errors;
  auto take = [&](int k, const char* v, const std::string& shown) {
    Opt& o = t[k];
    if (o.count >= o.maxcount) {
      errors.push_back("excess option " + shown);
      return;
    }
    if (o.kind != K_LIT) {
      std::string why;
      if (!store(cfg, k, o.count, v, why)) {
        errors.push_back("invalid argument \"" + std::string(v) + "\" to option " + label(o) +
                         "=" + o.datatype + (why.empty() ? "" : " (" + why + ")"));
        return;
      }
    }
    ++o.count;
  };
  bool only_args = false;
  for (int i = 1; i < argc; ++i) {
    const char* a = argv[i];
    if (!a) continue;
    if (only_args || a[0] != '-' || a[1] == 0) {
      errors.push_back("unexpected argument \"" + std::string(a) + "\"");
      continue;
    }
    if (!strcmp(a, "--")) {
      only_args = true;
      continue;
    }
    if (a[1] == '-') {
      std::string body(a + 2), val;
      const size_t eq = body.find('=');
      const bool has_eq = eq != std::string::npos;
      if (has_eq) val = body.substr(eq + 1), body = body.substr(0, eq);
      int hit = -1, nprefix = 0;
      for (int k = 0; k < O_COUNT && hit < 0; ++k)
        if (t[k].name && body == t[k].name) hit = k;
      if (hit < 0) {
        std::string first;
        for (int k = 0; k < O_COUNT; ++k)
          if (t[k].name && !body.empty() && strncmp(t[k].name, body.c_str(), body.size()) == 0 &&
              first != t[k].name) {
            if (hit < 0) hit = k, first = t[k].name;
            ++nprefix;
          }
        if (nprefix > 1) {
          errors.push_back("ambiguous option \"" + std::string(a) + "\"");
          continue;
        }
      }
      if (hit < 0) {
        errors.push_back("invalid option \"" + std::string(a) + "\"");
        continue;
      }
      const std::string shown = std::string("--") + t[hit].name;
      if (t[hit].kind == K_LIT) {
        if (has_eq) errors.push_back("option " + shown + " takes no value");
        else take(hit, nullptr, shown);
        continue;
      }
      if (!has_eq) {
        if (i + 1 >= argc || !argv[i + 1]) {
          errors.push_back("option " + shown + " requires an argument");
          continue;
        }
        val = argv[++i];
      }
      take(hit, val.c_str(), shown);
      continue;
    }
    // a cluster of short options; a value option takes the rest of the word or the next
    for (const char* p = a + 1; *p; ++p) {
      int hit = -1;
      for (int k = 0; k < O_COUNT && hit < 0; ++k)
        if (t[k].shortc == *p) hit = k;
      const std::string shown = std::string("-") + *p;
      if (hit < 0) {
        errors.push_back("invalid option " + shown);
        break;
      }
      if (t[hit].kind == K_LIT) {
        take(hit, nullptr, shown);
        continue;
      }
      const char* v = p[1] ? p + 1 : nullptr;
      if (!v) {
        if (i + 1 >= argc || !argv[i + 1]) {
          errors.push_back("option " + shown + " requires an argument");
          break;
        }
        v = argv[++i];
      }
      take(hit, v, shown);
      break;
    }
  }
  if (t[O_HELP].count > 0) {
    put(msg, msglen, help_text(prog, t));
    return WAP_CLI_EXIT;
  }
  if (t[O_VERSION].count > 0) {
    put(msg, msglen, std::string(prog) + ": MI355X warp-and-project driver, C ABI version " +
                         std::to_string(wap_version()) + "\n");
    return WAP_CLI_EXIT;
  }
  for (int k = 0; k < O_COUNT; ++k)
    if (t[k].count < t[k].mincount) errors.push_back("missing option " + label(t[k]));
  if (!errors.empty()) {
    std::string s;
    for (const std::string& e : errors) s += std::string(prog) + ": " + e + "\n";
    s += std::string("Try '") + prog + " --help' for more information.\n";
    put(msg, msglen, s);
    return WAP_E_ARG;
  }
  if (argc <= 1) {
    put(msg, msglen, std::string("Try '") + prog + " --help' for more information.\n");
    return WAP_CLI_EXIT;
  }
  return WAP_OK;
}
