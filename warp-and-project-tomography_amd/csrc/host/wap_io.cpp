// Host-side file formats of the warp-and-project loop: the projection-stack reader, the
// Analyze 7.5 writer and the angle list.
//
// The reference loads its projection stack with CImg's generic loader and splits it
// along z into one image per view (Project/WaP_Tomo.cpp:339), reads the view angles
// from a text file (Project/WaP_Tomo.cpp:324-331) and saves volumes, flow channels and
// warped volumes as Analyze 7.5 pairs through CImg::save_analyze (Project/CImg.h:51715,
// call sites Project/WaP_Tomo.cpp:766-771,1418,1669). These are plain readers/writers of
// the published formats; the Analyze header is filled field for field as that writer
// fills it so downstream tools see the same files.
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <strings.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <string>
#include <vector>

#include "../../../include/wap.h"

namespace {

thread_local std::string g_host_error;

int fail(int code, const std::string& msg) {
  g_host_error = msg;
  return code;
}

// Extension after the last '.' of the file-name part (lower case, "" if none).
std::string extension(const std::string& path) {
  const size_t slash = path.find_last_of('/');
  const size_t dot = path.find_last_of('.');
  if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return "";
  std::string e = path.substr(dot + 1);
  for (char& c : e) c = (char)tolower((unsigned char)c);
  return e;
}

bool read_file(const std::string& path, std::vector<unsigned char>& bytes) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return false;
  bytes.clear();
  unsigned char buf[1 << 16];
  size_t n;
  while ((n = fread(buf, 1, sizeof(buf), f)) > 0) bytes.insert(bytes.end(), buf, buf + n);
  const bool ok = !ferror(f);
  fclose(f);
  return ok;
}

template <class T>
T load_le(const unsigned char* p, bool swap) {
  unsigned char b[sizeof(T)];
  for (size_t i = 0; i < sizeof(T); ++i) b[i] = swap ? p[sizeof(T) - 1 - i] : p[i];
  T v;
  memcpy(&v, b, sizeof(T));
  return v;
}

// Element kinds shared by the readers.
enum Elem { E_U8, E_I8, E_U16, E_I16, E_U32, E_I32, E_F32, E_F64 };

size_t elem_size(Elem e) {
  switch (e) {
    case E_U8: case E_I8: return 1;
    case E_U16: case E_I16: return 2;
    case E_U32: case E_I32: case E_F32: return 4;
    default: return 8;
  }
}

// Converts n elements at src (byte-swapped if swap) to float, times scale when scale != 1
// (the product is taken in the element's promoted type, then rounded to float).
void convert(const unsigned char* src, size_t n, Elem e, bool swap, float scale, float* out) {
  const size_t s = elem_size(e);
  for (size_t i = 0; i < n; ++i) {
    const unsigned char* p = src + i * s;
    float v;
    switch (e) {
      case E_U8: v = (float)p[0] * scale; break;
      case E_I8: v = (float)(int8_t)p[0] * scale; break;
      case E_U16: v = (float)load_le<uint16_t>(p, swap) * scale; break;
      case E_I16: v = (float)load_le<int16_t>(p, swap) * scale; break;
      case E_U32: v = (float)load_le<uint32_t>(p, swap) * scale; break;
      case E_I32: v = (float)load_le<int32_t>(p, swap) * scale; break;
      case E_F32: v = load_le<float>(p, swap) * scale; break;
      default: v = (float)(load_le<double>(p, swap) * (double)scale); break;
    }
    out[i] = v;
  }
}

struct Image {
  int dims[4] = {1, 1, 1, 1};
  size_t count() const { return (size_t)dims[0] * dims[1] * dims[2] * dims[3]; }
};

int check_capacity(const Image& im, int* dims4, float* out, size_t capacity) {
  for (int a = 0; a < 4; ++a) dims4[a] = im.dims[a];
  if (out && capacity < im.count())
    return fail(WAP_E_ARG, "output buffer holds " + std::to_string(capacity) + " floats, image has " +
                               std::to_string(im.count()));
  return WAP_OK;
}

// ---------------------------------------------------------------- Analyze 7.5 / NIfTI-1
// Header fields used: sizeof_hdr @0 (348; a value >= 4096 means the other byte order),
// dim[0..4] @40 (shorts), datatype @70, vox_offset @108 (single-file form), scale @112
// (0 means 1). Datatypes 2 (u8), 4 (i16), 8 (i32), 16 (f32), 64 (f64).
int read_analyze(const std::string& path, const std::string& ext, int* dims4, float* out,
                 size_t capacity) {
  std::string hdr = path, img = path;
  const std::string stem = path.substr(0, path.size() - ext.size());
  if (ext == "hdr") img = stem + "img";
  if (ext == "img") hdr = stem + "hdr";
  std::vector<unsigned char> h;
  if (!read_file(hdr, h)) return fail(WAP_E_IO, "cannot read " + hdr);
  if (h.size() < 348) return fail(WAP_E_IO, hdr + ": short Analyze header");
  uint32_t hs = load_le<uint32_t>(h.data(), false);
  const bool swap = hs >= 4096;
  if (swap) hs = load_le<uint32_t>(h.data(), true);
  if (hs == 0 || hs > h.size()) return fail(WAP_E_IO, hdr + ": bad header size");
  const unsigned ndim = load_le<uint16_t>(&h[40], swap);
  Image im;
  for (unsigned a = 0; a < 4 && a < ndim; ++a) im.dims[a] = load_le<uint16_t>(&h[42 + 2 * a], swap);
  const int dt = load_le<int16_t>(&h[70], swap);
  float scale = load_le<float>(&h[112], swap);
  if (scale == 0.0f) scale = 1.0f;
  Elem e;
  switch (dt) {
    case 2: e = E_U8; break;
    case 4: e = E_I16; break;
    case 8: e = E_I32; break;
    case 16: e = E_F32; break;
    case 64: e = E_F64; break;
    default: return fail(WAP_E_IO, hdr + ": unsupported Analyze datatype " + std::to_string(dt));
  }
  int r = check_capacity(im, dims4, out, capacity);
  if (r != WAP_OK || !out) return r;
  std::vector<unsigned char> d;
  size_t off = 0;
  if (hdr == img) {
    d.swap(h);
    off = (size_t)load_le<float>(&d[108], swap);
  } else if (!read_file(img, d)) {
    return fail(WAP_E_IO, "cannot read " + img);
  }
  const size_t need = im.count() * elem_size(e);
  if (d.size() < off + need) return fail(WAP_E_IO, img + ": fewer voxels than the header states");
  convert(d.data() + off, im.count(), e, swap, scale, out);
  return WAP_OK;
}

// ---------------------------------------------------------------- TIFF (baseline)
// Uncompressed strips, one image per page (pages -> D), SamplesPerPixel interleaved
// samples (-> C), BitsPerSample 8/16/32/64 with SampleFormat 1 (unsigned), 2 (signed) or
// 3 (IEEE float), grey or RGB samples taken as stored. Tiled, compressed, palette and
// other colour-model files are refused.
struct TiffPage {
  uint32_t w = 0, h = 0, bits = 1, spp = 1, fmt = 1, comp = 1, planar = 1, rps = 0xffffffffu;
  uint32_t photometric = 1;
  std::vector<uint32_t> offs, counts;
  bool tiled = false;
};

int read_tiff(const std::string& path, int* dims4, float* out, size_t capacity) {
  std::vector<unsigned char> b;
  if (!read_file(path, b)) return fail(WAP_E_IO, "cannot read " + path);
  if (b.size() < 8) return fail(WAP_E_IO, path + ": not a TIFF file");
  bool swap;
  if (b[0] == 'I' && b[1] == 'I') swap = false;
  else if (b[0] == 'M' && b[1] == 'M') swap = true;
  else return fail(WAP_E_IO, path + ": not a TIFF file");
  if (load_le<uint16_t>(&b[2], swap) != 42) return fail(WAP_E_IO, path + ": not a classic TIFF file");
  auto in = [&](size_t off, size_t n) { return off <= b.size() && n <= b.size() - off; };
  std::vector<TiffPage> pages;
  uint32_t ifd = load_le<uint32_t>(&b[4], swap);
  while (ifd != 0) {
    if (!in(ifd, 2)) return fail(WAP_E_IO, path + ": bad IFD offset");
    if (pages.size() >= 65535) return fail(WAP_E_IO, path + ": too many pages (or an IFD cycle)");
    const unsigned n = load_le<uint16_t>(&b[ifd], swap);
    if (!in(ifd + 2, 12 * (size_t)n + 4)) return fail(WAP_E_IO, path + ": truncated IFD");
    TiffPage p;
    for (unsigned k = 0; k < n; ++k) {
      const unsigned char* e = &b[ifd + 2 + 12 * k];
      const unsigned tag = load_le<uint16_t>(e, swap), type = load_le<uint16_t>(e + 2, swap);
      const uint32_t cnt = load_le<uint32_t>(e + 4, swap);
      const size_t tsz = type == 3 ? 2 : (type == 4 ? 4 : (type == 1 ? 1 : 0));
      auto values = [&](std::vector<uint32_t>& v) -> bool {
        if (tsz == 0) return false;
        const size_t bytes = tsz * cnt;
        const unsigned char* src = e + 8;
        if (bytes > 4) {
          const uint32_t at = load_le<uint32_t>(e + 8, swap);
          if (!in(at, bytes)) return false;
          src = &b[at];
        }
        v.resize(cnt);
        for (uint32_t i = 0; i < cnt; ++i)
          v[i] = tsz == 2 ? load_le<uint16_t>(src + 2 * i, swap)
                          : (tsz == 4 ? load_le<uint32_t>(src + 4 * i, swap) : src[i]);
        return true;
      };
      std::vector<uint32_t> v;
      switch (tag) {
        case 256: case 257: case 258: case 259: case 262: case 277: case 278: case 284: case 339:
          if (!values(v) || v.empty()) return fail(WAP_E_IO, path + ": bad TIFF tag " + std::to_string(tag));
          if (tag == 256) p.w = v[0];
          if (tag == 257) p.h = v[0];
          if (tag == 258) p.bits = v[0];
          if (tag == 259) p.comp = v[0];
          if (tag == 262) p.photometric = v[0];
          if (tag == 277) p.spp = v[0];
          if (tag == 278) p.rps = v[0];
          if (tag == 284) p.planar = v[0];
          if (tag == 339) p.fmt = v[0];
          break;
        case 273:
          if (!values(p.offs)) return fail(WAP_E_IO, path + ": bad StripOffsets");
          break;
        case 279:
          if (!values(p.counts)) return fail(WAP_E_IO, path + ": bad StripByteCounts");
          break;
        case 322: case 323: case 324: case 325:
          p.tiled = true;
          break;
        default:
          break;
      }
    }
    pages.push_back(p);
    ifd = load_le<uint32_t>(&b[ifd + 2 + 12 * (size_t)n], swap);
  }
  if (pages.empty()) return fail(WAP_E_IO, path + ": no images");
  const TiffPage& p0 = pages[0];
  Elem e;
  if (p0.fmt == 3 && p0.bits == 32) e = E_F32;
  else if (p0.fmt == 3 && p0.bits == 64) e = E_F64;
  else if (p0.fmt == 2 && p0.bits == 8) e = E_I8;
  else if (p0.fmt == 2 && p0.bits == 16) e = E_I16;
  else if (p0.fmt == 2 && p0.bits == 32) e = E_I32;
  else if (p0.fmt == 1 && p0.bits == 8) e = E_U8;
  else if (p0.fmt == 1 && p0.bits == 16) e = E_U16;
  else if (p0.fmt == 1 && p0.bits == 32) e = E_U32;
  else return fail(WAP_E_IO, path + ": unsupported TIFF sample format");
  for (const TiffPage& p : pages) {
    if (p.tiled || p.comp != 1 || (p.planar != 1 && p.spp > 1))
      return fail(WAP_E_IO, path + ": only uncompressed, stripped, interleaved TIFF is read");
    if (p.photometric > 2)
      return fail(WAP_E_IO, path + ": only grey or RGB samples are read (photometric " +
                                std::to_string(p.photometric) + ")");
    if (p.w != p0.w || p.h != p0.h || p.bits != p0.bits || p.spp != p0.spp || p.fmt != p0.fmt)
      return fail(WAP_E_IO, path + ": pages differ in size or format");
    if (p.offs.empty() || p.offs.size() != p.counts.size())
      return fail(WAP_E_IO, path + ": missing strips");
  }
  if (p0.w == 0 || p0.h == 0 || p0.w > 65535 || p0.h > 65535 || p0.spp == 0 || p0.spp > 64 ||
      pages.size() > 65535)
    return fail(WAP_E_IO, path + ": image size out of range");
  Image im;
  im.dims[0] = (int)p0.w;
  im.dims[1] = (int)p0.h;
  im.dims[2] = (int)pages.size();
  im.dims[3] = (int)p0.spp;
  int r = check_capacity(im, dims4, out, capacity);
  if (r != WAP_OK || !out) return r;
  const size_t es = elem_size(e), plane = (size_t)p0.w * p0.h;
  const size_t page_bytes = plane * p0.spp * es;
  std::vector<unsigned char> raw(page_bytes);
  std::vector<float> vals(plane * p0.spp);
  for (size_t z = 0; z < pages.size(); ++z) {
    const TiffPage& p = pages[z];
    size_t got = 0;
    for (size_t s = 0; s < p.offs.size() && got < page_bytes; ++s) {
      const size_t n = std::min<size_t>(p.counts[s], page_bytes - got);
      if (!in(p.offs[s], n)) return fail(WAP_E_IO, path + ": strip outside the file");
      memcpy(raw.data() + got, &b[p.offs[s]], n);
      got += n;
    }
    if (got < page_bytes) return fail(WAP_E_IO, path + ": strips hold fewer samples than the page");
    convert(raw.data(), plane * p0.spp, e, swap, 1.0f, vals.data());
    const size_t vol = plane * pages.size();
    for (size_t px = 0; px < plane; ++px)
      for (uint32_t c = 0; c < p0.spp; ++c) out[c * vol + z * plane + px] = vals[px * p0.spp + c];
  }
  return WAP_OK;
}

// ---------------------------------------------------------------- MetaImage (.mha/.mhd)
int read_metaimage(const std::string& path, int* dims4, float* out, size_t capacity) {
  std::vector<unsigned char> b;
  if (!read_file(path, b)) return fail(WAP_E_IO, "cannot read " + path);
  size_t pos = 0, data_off = 0;
  std::string datafile, etype;
  bool msb = false, compressed = false;
  long header_size = -1;
  int ndims = 0, nch = 1;
  std::vector<long> dsz;
  while (pos < b.size()) {
    size_t eol = pos;
    while (eol < b.size() && b[eol] != '\n') ++eol;
    std::string line(b.begin() + pos, b.begin() + eol);
    pos = eol < b.size() ? eol + 1 : eol;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    auto trim = [](std::string s) {
      const size_t a = s.find_first_not_of(" \t"), z = s.find_last_not_of(" \t");
      return a == std::string::npos ? std::string() : s.substr(a, z - a + 1);
    };
    const std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
    if (key == "NDims") ndims = atoi(val.c_str());
    else if (key == "DimSize") {
      dsz.clear();
      char* s = const_cast<char*>(val.c_str());
      while (*s) {
        char* end;
        const long v = strtol(s, &end, 10);
        if (end == s) break;
        dsz.push_back(v);
        s = end;
      }
    } else if (key == "ElementType") etype = val;
    else if (key == "ElementNumberOfChannels") nch = atoi(val.c_str());
    else if (key == "HeaderSize") header_size = atol(val.c_str());
    else if (key == "BinaryDataByteOrderMSB" || key == "ElementByteOrderMSB")
      msb = strcasecmp(val.c_str(), "true") == 0;
    else if (key == "CompressedData") compressed = strcasecmp(val.c_str(), "true") == 0;
    else if (key == "ElementDataFile") {
      datafile = val;
      data_off = pos;
      break;
    }
  }
  if (datafile.empty() || ndims < 1 || ndims > 4 || (int)dsz.size() < ndims)
    return fail(WAP_E_IO, path + ": incomplete MetaImage header");
  if (compressed) return fail(WAP_E_IO, path + ": compressed MetaImage data is not read");
  Elem e;
  if (etype == "MET_UCHAR") e = E_U8;
  else if (etype == "MET_CHAR") e = E_I8;
  else if (etype == "MET_USHORT") e = E_U16;
  else if (etype == "MET_SHORT") e = E_I16;
  else if (etype == "MET_UINT") e = E_U32;
  else if (etype == "MET_INT") e = E_I32;
  else if (etype == "MET_FLOAT") e = E_F32;
  else if (etype == "MET_DOUBLE") e = E_F64;
  else return fail(WAP_E_IO, path + ": unsupported ElementType " + etype);
  Image im;
  for (int a = 0; a < ndims && a < 3; ++a) {
    if (dsz[a] <= 0 || dsz[a] > (1L << 20)) return fail(WAP_E_IO, path + ": bad DimSize");
    im.dims[a] = (int)dsz[a];
  }
  if (ndims == 4) {
    if (nch != 1) return fail(WAP_E_IO, path + ": 4-D data with channels is not read");
    nch = (int)dsz[3];
  }
  if (nch < 1 || nch > 64) return fail(WAP_E_IO, path + ": bad channel count");
  im.dims[3] = nch;
  int r = check_capacity(im, dims4, out, capacity);
  if (r != WAP_OK || !out) return r;
  std::vector<unsigned char> ext;
  const unsigned char* src;
  size_t avail;
  if (datafile == "LOCAL") {
    src = b.data() + data_off;
    avail = b.size() - data_off;
  } else {
    const size_t slash = path.find_last_of('/');
    const std::string dir = slash == std::string::npos ? "" : path.substr(0, slash + 1);
    const std::string df = datafile[0] == '/' ? datafile : dir + datafile;
    if (!read_file(df, ext)) return fail(WAP_E_IO, "cannot read " + df);
    src = ext.data();
    avail = ext.size();
  }
  const size_t need = im.count() * elem_size(e);
  if (header_size == -1 && datafile != "LOCAL" && avail >= need) {
    // HeaderSize = -1: the data are the last bytes of the file
    src += avail - need;
    avail = need;
  } else if (header_size > 0) {
    if ((size_t)header_size > avail) return fail(WAP_E_IO, path + ": HeaderSize beyond the data");
    src += header_size;
    avail -= header_size;
  }
  if (avail < need) return fail(WAP_E_IO, path + ": fewer voxels than DimSize states");
  // MetaImage stores channels interleaved per voxel
  std::vector<float> vals(im.count());
  convert(src, im.count(), e, msb, 1.0f, vals.data());
  const size_t vox = im.count() / nch;
  for (size_t v = 0; v < vox; ++v)
    for (int c = 0; c < nch; ++c) out[c * vox + v] = vals[v * nch + c];
  return WAP_OK;
}

// ---------------------------------------------------------------- raw float32
int read_raw(const std::string& path, int* dims4, float* out, size_t capacity) {
  Image im;
  for (int a = 0; a < 4; ++a) {
    const int d = a == 3 && dims4[a] == 0 ? 1 : dims4[a];
    if (d <= 0) return fail(WAP_E_ARG, path + ": a raw float32 file needs its dimensions");
    im.dims[a] = d;
  }
  struct stat st;
  if (stat(path.c_str(), &st) != 0) return fail(WAP_E_IO, "cannot read " + path);
  if ((size_t)st.st_size != im.count() * sizeof(float))
    return fail(WAP_E_IO, path + ": size " + std::to_string((long long)st.st_size) +
                              " bytes, dimensions need " + std::to_string(im.count() * 4));
  int r = check_capacity(im, dims4, out, capacity);
  if (r != WAP_OK || !out) return r;
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return fail(WAP_E_IO, "cannot read " + path);
  const size_t n = fread(out, sizeof(float), im.count(), f);
  fclose(f);
  if (n != im.count()) return fail(WAP_E_IO, path + ": short read");
  return WAP_OK;
}

bool write_all(const std::string& path, const void* a, size_t na, const void* b, size_t nb) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return false;
  bool ok = fwrite(a, 1, na, f) == na;
  if (ok && nb) ok = fwrite(b, 1, nb, f) == nb;
  ok = (fclose(f) == 0) && ok;
  return ok;
}

}  // namespace

extern "C" const char* wap_host_error(void) { return g_host_error.c_str(); }

extern "C" int wap_read_image(const char* path, int* dims4, float* out, size_t capacity) {
  if (!path || !dims4) return fail(WAP_E_ARG, "wap_read_image: null path or dims");
  const std::string p(path), e = extension(p);
  if (e == "hdr" || e == "img" || e == "nii") return read_analyze(p, e, dims4, out, capacity);
  if (e == "tif" || e == "tiff") return read_tiff(p, dims4, out, capacity);
  if (e == "mha" || e == "mhd") return read_metaimage(p, dims4, out, capacity);
  return read_raw(p, dims4, out, capacity);
}

extern "C" int wap_write_analyze(const char* path, const int* dims4, const float* data,
                                 const float* voxel_size) {
  if (!path || !dims4) return fail(WAP_E_ARG, "wap_write_analyze: null path or dims");
  size_t n = 1;
  for (int a = 0; a < 4; ++a) {
    if (dims4[a] < 0 || dims4[a] > 32767)
      return fail(WAP_E_ARG, "wap_write_analyze: dimension out of the Analyze range");
    n *= (size_t)dims4[a];
  }
  const std::string p(path), e = extension(p);
  std::string hname, iname;
  if (e.empty()) {
    hname = p + ".hdr";
    iname = p + ".img";
  } else if (e == "hdr" || e == "img") {
    const std::string stem = p.substr(0, p.size() - 3);
    hname = stem + "hdr";
    iname = stem + "img";
  } else if (e == "nii") {
    hname = p;
  } else {
    return fail(WAP_E_ARG, "wap_write_analyze: " + p + " is not .hdr, .img or .nii");
  }
  if (n == 0) {
    // an empty image leaves an empty file under the given name
    if (!write_all(p, nullptr, 0, nullptr, 0)) return fail(WAP_E_IO, "cannot write " + p);
    return WAP_OK;
  }
  if (!data) return fail(WAP_E_ARG, "wap_write_analyze: null data");
  const bool single = iname.empty();
  std::vector<unsigned char> h(single ? 352 : 348, 0);
  auto put_i32 = [&](size_t at, int32_t v) { memcpy(&h[at], &v, 4); };
  auto put_i16 = [&](size_t at, int16_t v) { memcpy(&h[at], &v, 2); };
  auto put_f32 = [&](size_t at, float v) { memcpy(&h[at], &v, 4); };
  put_i32(0, 348);                   // sizeof_hdr
  memcpy(&h[4], "CImg", 5);          // data_type
  memcpy(&h[14], " ", 2);            // db_name
  put_i16(36, 4096);                 // session_error
  h[38] = 114;                       // regular = 'r'
  put_i16(40, 4);                    // dim[0]
  for (int a = 0; a < 4; ++a) put_i16(42 + 2 * a, (int16_t)dims4[a]);
  put_i16(70, 16);                   // datatype: float
  put_i16(72, (int16_t)sizeof(float));  // bitpix field as that writer fills it (bytes)
  put_f32(76, 0.0f);                 // pixdim[0]
  for (int a = 0; a < 3; ++a) put_f32(80 + 4 * a, voxel_size ? voxel_size[a] : 1.0f);
  put_f32(108, single ? 352.0f : 0.0f);  // vox_offset
  put_f32(112, 1.0f);                // scale factor
  const size_t nb = n * sizeof(float);
  if (single) {
    if (!write_all(hname, h.data(), h.size(), data, nb)) return fail(WAP_E_IO, "cannot write " + hname);
  } else {
    if (!write_all(hname, h.data(), h.size(), nullptr, 0)) return fail(WAP_E_IO, "cannot write " + hname);
    if (!write_all(iname, data, nb, nullptr, 0)) return fail(WAP_E_IO, "cannot write " + iname);
  }
  return WAP_OK;
}

extern "C" int wap_read_angles(const char* path, int n, double* out, int* nread) {
  if (!path || n < 0 || (n > 0 && !out)) return fail(WAP_E_ARG, "wap_read_angles: bad arguments");
  for (int i = 0; i < n; ++i) out[i] = 0.0;
  if (nread) *nread = 0;
  std::ifstream in(path);
  if (!in) return fail(WAP_E_IO, std::string("cannot read ") + path);
  int k = 0;
  while (k < n && (in >> out[k])) ++k;
  if (k < n) out[k] = 0.0;  // a failed extraction may have stored into the slot
  if (nread) *nread = k;
  return WAP_OK;
}
