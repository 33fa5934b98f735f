// imaging.hip -- host-side data path of oap's PCA application: PNG images
// -> cropped pixel vectors -> the data matrix (one image per column) that
// RecToSquareApi squares and the eigensolver decomposes.  No kernels; the
// matrix is copied to HBM once.
//
// Reference:
//   oapAppUtils/PngFile.cpp:107-184,380-470  (raw 8-bit rows, colour count =
//       row bytes / width, crop to the optimal window, pixel = r<<16|g<<8|b)
//   oapAppUtils/GraphicsUtilsImpl.h         (GetOptWidth / GetOptHeight: the
//       bounding box of the byte columns / pixel rows that are not constant)
//   oapAppUtils/Image.cpp:119-178           (float = pixel / 0xFFFFFF)
//   oapAppUtils/DataLoader.cpp:30-250       (one window size for all images:
//       the load process with forced output sizes and reloads; one column
//       per image)
// PNG decoding is this file's own (zlib inflate, the five scanline filters);
// the reference reads PNGs through libpng.
#include <zlib.h>

#include <cstdio>
#include <cstring>
#include <functional>
#include <new>
#include <string>
#include <vector>

#include "internal.h"

namespace oaph {
namespace {

// ---------------------------------------------------------------- PNG
struct Png {
  uint32_t width = 0, height = 0, colors = 0;
  std::vector<uint8_t> rows;  // height x width*colors bytes
};

uint32_t be32(const uint8_t* p) {
  return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}

void put32(std::vector<uint8_t>& v, uint32_t x) {
  v.push_back(x >> 24);
  v.push_back(x >> 16);
  v.push_back(x >> 8);
  v.push_back(x);
}

uint8_t paeth(int a, int b, int c) {
  const int p = a + b - c;
  const int pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
  if (pa <= pb && pa <= pc) return (uint8_t)a;
  if (pb <= pc) return (uint8_t)b;
  return (uint8_t)c;
}

// 8-bit, non-interlaced PNGs of colour type 0 (grey), 2 (RGB), 3 (palette
// indices, one byte as libpng delivers them untransformed), 4 (grey+alpha),
// 6 (RGBA)
oaph_status png_decode(const std::vector<uint8_t>& f, Png* out) {
  static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  OAPH_CHECK(f.size() >= 8 && std::memcmp(f.data(), sig, 8) == 0, OAPH_ERR_ARG);
  size_t o = 8;
  std::vector<uint8_t> idat;
  bool ihdr = false;
  uint32_t w = 0, h = 0, bpp = 0;
  while (o + 12 <= f.size()) {
    const uint32_t len = be32(&f[o]);
    OAPH_CHECK(o + 12 + (size_t)len <= f.size(), OAPH_ERR_ARG);
    const uint8_t* type = &f[o + 4];
    const uint8_t* data = &f[o + 8];
    const uint32_t crc = be32(&f[o + 8 + len]);
    OAPH_CHECK(crc32(crc32(0L, Z_NULL, 0), type, len + 4) == crc, OAPH_ERR_ARG);
    if (!std::memcmp(type, "IHDR", 4)) {
      OAPH_CHECK(len == 13, OAPH_ERR_ARG);
      w = be32(data);
      h = be32(data + 4);
      const uint8_t depth = data[8], ctype = data[9], interlace = data[12];
      OAPH_CHECK(depth == 8 && interlace == 0 && data[10] == 0 && data[11] == 0,
                 OAPH_ERR_UNSUPPORTED);
      switch (ctype) {
        case 0: bpp = 1; break;
        case 2: bpp = 3; break;
        case 3: bpp = 1; break;
        case 4: bpp = 2; break;
        case 6: bpp = 4; break;
        default: return OAPH_ERR_UNSUPPORTED;
      }
      OAPH_CHECK(w > 0 && h > 0 && (size_t)w * h * bpp < ((size_t)1 << 32), OAPH_ERR_UNSUPPORTED);
      ihdr = true;
    } else if (!std::memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), data, data + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      break;
    }
    o += 12 + (size_t)len;
  }
  OAPH_CHECK(ihdr && !idat.empty(), OAPH_ERR_ARG);
  const size_t stride = (size_t)w * bpp;
  std::vector<uint8_t> raw((stride + 1) * h);
  uLongf rawlen = (uLongf)raw.size();
  OAPH_CHECK(uncompress(raw.data(), &rawlen, idat.data(), (uLong)idat.size()) == Z_OK &&
                 rawlen == raw.size(),
             OAPH_ERR_ARG);
  out->width = w;
  out->height = h;
  out->colors = bpp;
  out->rows.assign(stride * h, 0);
  for (uint32_t y = 0; y < h; ++y) {
    const uint8_t ft = raw[y * (stride + 1)];
    const uint8_t* in = &raw[y * (stride + 1) + 1];
    uint8_t* cur = &out->rows[y * stride];
    const uint8_t* prev = y ? &out->rows[(y - 1) * stride] : nullptr;
    for (size_t x = 0; x < stride; ++x) {
      const int a = x >= bpp ? cur[x - bpp] : 0, b = prev ? prev[x] : 0,
                c = (prev && x >= bpp) ? prev[x - bpp] : 0;
      int v = in[x];
      switch (ft) {
        case 0: break;
        case 1: v += a; break;
        case 2: v += b; break;
        case 3: v += (a + b) >> 1; break;
        case 4: v += paeth(a, b, c); break;
        default: return OAPH_ERR_ARG;
      }
      cur[x] = (uint8_t)v;
    }
  }
  return OAPH_OK;
}

oaph_status read_file(const char* path, std::vector<uint8_t>* out) {
  FILE* fp = std::fopen(path, "rb");
  OAPH_CHECK(fp, OAPH_ERR_ARG);
  std::fseek(fp, 0, SEEK_END);
  const long n = std::ftell(fp);
  std::fseek(fp, 0, SEEK_SET);
  out->resize(n > 0 ? (size_t)n : 0);
  const bool ok = n > 0 && std::fread(out->data(), 1, (size_t)n, fp) == (size_t)n;
  std::fclose(fp);
  return ok ? OAPH_OK : OAPH_ERR_ARG;
}

// ---------------------------------------------------------------- optimal window
// GetOptWidth: the first and last BYTE column (of width * colors) that is not
// constant down the rows, as pixel columns (byte / colors); the whole width
// when every column is constant.  GetOptHeight: the first and last pixel row
// whose pixels are not all equal to its first pixel.
struct Window {
  uint32_t size = 0, begin = 0;
};

bool column_constant(const uint8_t* bm, size_t stride, uint32_t h, size_t byte) {
  const uint8_t g = bm[byte];
  for (uint32_t y = 0; y < h; ++y)
    if (bm[(size_t)y * stride + byte] != g) return false;
  return true;
}

bool row_constant(const uint8_t* bm, size_t stride, uint32_t w, uint32_t colors, uint32_t y) {
  const uint8_t* r = bm + (size_t)y * stride;
  for (uint32_t x = 0; x < w; ++x)
    if (std::memcmp(r + (size_t)x * colors, r, colors) != 0) return false;
  return true;
}

Window opt_width(const uint8_t* bm, uint32_t w, uint32_t h, uint32_t colors) {
  const size_t stride = (size_t)w * colors;
  size_t b = 0;
  while (b < stride && column_constant(bm, stride, h, b)) ++b;
  if (b == stride) return {w, 0};
  const uint32_t begin = (uint32_t)(b / colors);
  // the reference scans down to byte column 1 and yields 0 when none differs
  uint32_t end = 0;
  for (size_t e = stride - 1; e >= 1; --e)
    if (!column_constant(bm, stride, h, e)) {
      end = (uint32_t)(e / colors) + 1;
      break;
    }
  return {end - begin, begin};
}

Window opt_height(const uint8_t* bm, uint32_t w, uint32_t h, uint32_t colors) {
  const size_t stride = (size_t)w * colors;
  uint32_t b = 0;
  while (b < h && row_constant(bm, stride, w, colors, b)) ++b;
  if (b == h) return {h, 0};
  uint32_t end = 0;
  for (uint32_t e = h - 1; e >= 1; --e)
    if (!row_constant(bm, stride, w, colors, e)) {
      end = e + 1;
      break;
    }
  return {end - b, b};
}

// ---------------------------------------------------------------- load process
// DataLoader::executeLoadProcess over the images' own windows: the first
// image sets the reference window per dimension; a smaller image is forced
// to the reference window (and reloaded); a larger one becomes the reference
// and every earlier image is forced to both reference windows and run through
// the process again.  `loads` counts the bitmap loads the reference performs.
struct Plan {
  std::vector<Window> own_w, own_h, out_w, out_h;
  std::vector<bool> forced_w, forced_h;
  std::vector<uint32_t> loads;
};

Window current(const std::vector<Window>& own, const std::vector<Window>& out,
               const std::vector<bool>& forced, size_t i) {
  return forced[i] ? out[i] : own[i];
}

void load_process(Plan& P, Window refw, Window refh, size_t begin, size_t end) {
  for (size_t fa = begin; fa < end; ++fa) {
    ++P.loads[fa];
    const Window iw = current(P.own_w, P.out_w, P.forced_w, fa);
    const Window ih = current(P.own_h, P.out_h, P.forced_h, fa);
    bool reload = false, previous = false;
    auto verify = [&](Window& ref, Window img, std::vector<Window>& out, std::vector<bool>& forced) {
      if (ref.size == 0) {
        ref = img;
      } else if (img.size < ref.size) {
        out[fa] = ref;
        forced[fa] = true;
        reload = true;
      } else if (img.size > ref.size) {
        previous = true;
        ref = img;
      }
    };
    verify(refw, iw, P.out_w, P.forced_w);
    verify(refh, ih, P.out_h, P.forced_h);
    if (reload) ++P.loads[fa];
    if (previous) {
      for (size_t q = begin; q < fa; ++q) {
        P.out_w[q] = refw;
        P.out_h[q] = refh;
        P.forced_w[q] = P.forced_h[q] = true;
      }
      load_process(P, refw, refh, begin, fa);
    }
  }
}

void make_plan(Plan& P) {
  const size_t n = P.own_w.size();
  P.out_w = P.own_w;
  P.out_h = P.own_h;
  P.forced_w.assign(n, false);
  P.forced_h.assign(n, false);
  P.loads.assign(n, 0);
  load_process(P, Window{}, Window{}, 0, n);
  for (size_t i = 0; i < n; ++i) {
    if (!P.forced_w[i]) P.out_w[i] = P.own_w[i];
    if (!P.forced_h[i]) P.out_h[i] = P.own_h[i];
  }
}

}  // namespace
}  // namespace oaph

// PCA data loader: decoded images, their windows and the resulting columns
struct oaph_dataloader {
  std::vector<oaph::Png> images;
  oaph::Plan plan;
  uint32_t length = 0;  // pixels per column (common window area)
};

using namespace oaph;

namespace {

// pixel = r<<16 | g<<8 | b (grey: p, p, p; RGBA: alpha ignored) over the
// window rows and columns (row-major); the column value is pixel / 0xFFFFFF
oaph_status image_pixels(const oaph_dataloader* d, uint32_t i, uint32_t* out) {
  const Png& im = d->images[i];
  const Window W = d->plan.out_w[i], H = d->plan.out_h[i];
  OAPH_CHECK(W.begin + W.size <= im.width && H.begin + H.size <= im.height, OAPH_ERR_SHAPE);
  OAPH_CHECK(im.colors == 1 || im.colors == 3 || im.colors == 4, OAPH_ERR_UNSUPPORTED);
  const size_t stride = (size_t)im.width * im.colors;
  size_t k = 0;
  for (uint32_t y = H.begin; y < H.begin + H.size; ++y) {
    const uint8_t* row = &im.rows[(size_t)y * stride];
    for (uint32_t x = W.begin; x < W.begin + W.size; ++x) {
      const uint8_t* p = row + (size_t)x * im.colors;
      const uint32_t r = p[0], g = im.colors == 1 ? p[0] : p[1], b = im.colors == 1 ? p[0] : p[2];
      out[k++] = r << 16 | g << 8 | b;
    }
  }
  return OAPH_OK;
}

constexpr double kMaxPixel = 16777215.0;

oaph_status image_column(const oaph_dataloader* d, uint32_t i, double* out) {
  std::vector<uint32_t> px(d->length);
  OAPH_TRY(image_pixels(d, i, px.data()));
  for (uint32_t k = 0; k < d->length; ++k) out[k] = (double)px[k] / kMaxPixel;
  return OAPH_OK;
}

// value = pixel / 0xFFFFFF (correctly rounded division, the host's bits)
__global__ __launch_bounds__(256) void k_pixel_values(const uint32_t* __restrict__ px,
                                                      double* __restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = (double)px[i] / kMaxPixel;
}

}  // namespace

extern "C" {

oaph_status oaph_png_read(const char* path, uint32_t* width, uint32_t* height, uint32_t* colors,
                          uint8_t* pixels, size_t cap) {
  OAPH_CHECK(path && width && height && colors, OAPH_ERR_ARG);
  std::vector<uint8_t> f;
  OAPH_TRY(read_file(path, &f));
  Png p;
  OAPH_TRY(png_decode(f, &p));
  *width = p.width;
  *height = p.height;
  *colors = p.colors;
  if (pixels) {
    OAPH_CHECK(cap >= p.rows.size(), OAPH_ERR_ARG);
    std::memcpy(pixels, p.rows.data(), p.rows.size());
  }
  return OAPH_OK;
}

oaph_status oaph_png_write(const char* path, uint32_t width, uint32_t height, uint32_t colors,
                           const uint8_t* pixels) {
  OAPH_CHECK(path && pixels && width > 0 && height > 0, OAPH_ERR_ARG);
  uint8_t ctype = 0;
  switch (colors) {
    case 1: ctype = 0; break;
    case 2: ctype = 4; break;
    case 3: ctype = 2; break;
    case 4: ctype = 6; break;
    default: return OAPH_ERR_ARG;
  }
  const size_t stride = (size_t)width * colors;
  std::vector<uint8_t> raw((stride + 1) * height);
  for (uint32_t y = 0; y < height; ++y) {
    raw[y * (stride + 1)] = 0;  // filter: none
    std::memcpy(&raw[y * (stride + 1) + 1], pixels + (size_t)y * stride, stride);
  }
  uLongf zlen = compressBound((uLong)raw.size());
  std::vector<uint8_t> z(zlen);
  OAPH_CHECK(compress(z.data(), &zlen, raw.data(), (uLong)raw.size()) == Z_OK, OAPH_ERR_ARG);
  z.resize(zlen);
  std::vector<uint8_t> f = {137, 80, 78, 71, 13, 10, 26, 10};
  auto chunk = [&](const char* type, const std::vector<uint8_t>& data) {
    put32(f, (uint32_t)data.size());
    const size_t t0 = f.size();
    f.insert(f.end(), type, type + 4);
    f.insert(f.end(), data.begin(), data.end());
    put32(f, (uint32_t)crc32(crc32(0L, Z_NULL, 0), &f[t0], (uInt)(data.size() + 4)));
  };
  std::vector<uint8_t> ihdr;
  put32(ihdr, width);
  put32(ihdr, height);
  ihdr.insert(ihdr.end(), {8, ctype, 0, 0, 0});
  chunk("IHDR", ihdr);
  chunk("IDAT", z);
  chunk("IEND", {});
  FILE* fp = std::fopen(path, "wb");
  OAPH_CHECK(fp, OAPH_ERR_ARG);
  bool ok = std::fwrite(f.data(), 1, f.size(), fp) == f.size();
  ok = std::fclose(fp) == 0 && ok;
  return ok ? OAPH_OK : OAPH_ERR_ARG;
}

oaph_status oaph_image_opt_size(const uint8_t* bitmap, uint32_t width, uint32_t height,
                                uint32_t colors, uint32_t out[4]) {
  OAPH_CHECK(bitmap && out && width > 0 && height > 0 && colors > 0, OAPH_ERR_ARG);
  const Window w = opt_width(bitmap, width, height, colors);
  const Window h = opt_height(bitmap, width, height, colors);
  out[0] = w.size;
  out[1] = w.begin;
  out[2] = h.size;
  out[3] = h.begin;
  return OAPH_OK;
}

oaph_status oaph_dataloader_plan(const uint32_t* sizes, uint32_t count, uint32_t* out,
                                 uint32_t* loads) {
  OAPH_CHECK(sizes && out && count > 0, OAPH_ERR_ARG);
  Plan P;
  for (uint32_t i = 0; i < count; ++i) {
    P.own_w.push_back({sizes[4 * i], sizes[4 * i + 1]});
    P.own_h.push_back({sizes[4 * i + 2], sizes[4 * i + 3]});
  }
  make_plan(P);
  for (uint32_t i = 0; i < count; ++i) {
    out[4 * i] = P.out_w[i].size;
    out[4 * i + 1] = P.out_w[i].begin;
    out[4 * i + 2] = P.out_h[i].size;
    out[4 * i + 3] = P.out_h[i].begin;
    if (loads) loads[i] = P.loads[i];
  }
  return OAPH_OK;
}

oaph_status oaph_dataloader_create(oaph_dataloader** out, const char* const* paths,
                                   uint32_t count, int truncate) {
  OAPH_CHECK(out && paths && count > 0, OAPH_ERR_ARG);
  *out = nullptr;
  oaph_dataloader* d = new (std::nothrow) oaph_dataloader;
  OAPH_CHECK(d, OAPH_ERR_NOMEM);
  d->images.resize(count);
  for (uint32_t i = 0; i < count; ++i) {
    std::vector<uint8_t> f;
    oaph_status s = paths[i] ? read_file(paths[i], &f) : OAPH_ERR_ARG;
    if (s == OAPH_OK) s = png_decode(f, &d->images[i]);
    if (s != OAPH_OK) {
      delete d;
      return s;
    }
    const Png& im = d->images[i];
    const Window w = truncate ? opt_width(im.rows.data(), im.width, im.height, im.colors)
                              : Window{im.width, 0};
    const Window h = truncate ? opt_height(im.rows.data(), im.width, im.height, im.colors)
                              : Window{im.height, 0};
    d->plan.own_w.push_back(w);
    d->plan.own_h.push_back(h);
  }
  make_plan(d->plan);
  d->length = d->plan.out_w[0].size * d->plan.out_h[0].size;
  for (uint32_t i = 0; i < count; ++i) {
    const Png& im = d->images[i];
    const Window W = d->plan.out_w[i], H = d->plan.out_h[i];
    // every column has the same length (DataLoader::loadColumnVector throws
    // NotIdenticalLengths otherwise) and the window lies inside the image
    const bool ok = W.size * H.size == d->length && W.begin + W.size <= im.width &&
                    H.begin + H.size <= im.height &&
                    (im.colors == 1 || im.colors == 3 || im.colors == 4);
    if (!ok) {
      delete d;
      return OAPH_ERR_SHAPE;
    }
  }
  *out = d;
  return OAPH_OK;
}

oaph_status oaph_dataloader_destroy(oaph_dataloader* d) {
  delete d;
  return OAPH_OK;
}

oaph_status oaph_dataloader_info(const oaph_dataloader* d, oaph_matrix_info* info) {
  OAPH_CHECK(d && info, OAPH_ERR_ARG);
  info->isRe = 1;
  info->isIm = 0;
  info->columns = (uint32_t)d->images.size();
  info->rows = d->length;
  return OAPH_OK;
}

oaph_status oaph_dataloader_window(const oaph_dataloader* d, uint32_t image, uint32_t out[4],
                                   uint32_t* loads) {
  OAPH_CHECK(d && out && image < d->images.size(), OAPH_ERR_ARG);
  out[0] = d->plan.out_w[image].size;
  out[1] = d->plan.out_w[image].begin;
  out[2] = d->plan.out_h[image].size;
  out[3] = d->plan.out_h[image].begin;
  if (loads) *loads = d->plan.loads[image];
  return OAPH_OK;
}

oaph_status oaph_dataloader_column(const oaph_dataloader* d, uint32_t image, double* out) {
  OAPH_CHECK(d && out && image < d->images.size(), OAPH_ERR_ARG);
  return image_column(d, image, out);
}

// M (one image per column): each image's packed pixels are gathered on the
// host, contiguous per image, uploaded once (4 bytes a pixel), converted to
// values and transposed into place on the device (a host scatter at a
// stride of one row of M per pixel was most of the assembly time).
oaph_status oaph_dataloader_matrix(const oaph_dataloader* d, oaph_matrix* m) {
  OAPH_CHECK(d && m && m->reValues, OAPH_ERR_ARG);
  const uint32_t cols = (uint32_t)d->images.size();
  OAPH_CHECK(m->columns == cols && m->rows == d->length, OAPH_ERR_SHAPE);
  const size_t n = (size_t)cols * d->length;
  if ((d->length + 31) / 32 > 65535) {  // beyond the transpose grid: scatter on the host
    std::vector<double> rows(n), col(d->length);
    for (uint32_t i = 0; i < cols; ++i) {
      OAPH_TRY(image_column(d, i, col.data()));
      for (uint32_t r = 0; r < d->length; ++r) rows[(size_t)r * cols + i] = col[r];
    }
    return oaph_matrix_copy_h2d(m, rows.data(), nullptr);
  }
  std::vector<uint32_t> host(n);
  for (uint32_t i = 0; i < cols; ++i) OAPH_TRY(image_pixels(d, i, host.data() + (size_t)i * d->length));
  OAPH_TRY(ensure_init());
  DevBuf px, vals;
  oaph_status st = px.reserve(n * sizeof(uint32_t));
  if (st == OAPH_OK) st = vals.reserve(n * sizeof(double));
  if (st == OAPH_OK &&
      hipMemcpy(px.ptr, host.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
    st = OAPH_ERR_HIP;
  if (st == OAPH_OK) {
    const size_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_pixel_values, dim3(blocks < 65536 ? (unsigned)blocks : 65536u), dim3(256),
                       0, nullptr, px.as<uint32_t>(), vals.as<double>(), n);
    if (hipGetLastError() != hipSuccess) st = OAPH_ERR_HIP;
  }
  if (st == OAPH_OK) {
    // images as the rows of a cols x length matrix; M is its transpose
    oaph_matrix a{d->length, cols, d->length, cols, vals.as<double>(), nullptr};
    oaph_matrix out = *m;
    out.imValues = nullptr;
    st = launch_transpose(nullptr, &out, &a, false);
    if (st == OAPH_OK && hipStreamSynchronize(nullptr) != hipSuccess) st = OAPH_ERR_HIP;
  }
  px.release();
  vals.release();
  return st;
}

}  // extern "C"
