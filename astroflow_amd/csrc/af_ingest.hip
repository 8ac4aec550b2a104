// af_ingest.hip -- file readers that feed the plans and the server (host
// code).  Rows come out time-major, channels in ascending frequency, as the
// plans take them, and can be read straight into page-locked buffers
// (af_host_alloc), whose uploads to the device are asynchronous.
//
// SIGPROC filterbank (replaces Filterbank::read_header / the data read and
// the band flip of src/data/filterbank.cpp:319-472,793-842): the keyword
// header is parsed once; af_fil_read then reads any range of rows with
// positioned reads (pread), split over threads, and reverses the channel
// order of each row when the file's band descends (foff < 0), the header
// then describing the ascending band (fch1 + (nchans-1) foff, -foff).
//
// PSRFITS search mode (replaces PsrFits._load_data,
// python/astroflow/io/psrfits.py:32-76, without astropy): the SUBINT binary
// table's DATA column holds [NSBLK][NPOL][NCHAN] samples per row; pol 0, or
// the mean of pols 0 and 1 for NPOL >= 2, is returned.  fch1 = OBSFREQ -
// OBSBW/2, foff = CHAN_BW, tsamp = TBIN; a negative CHAN_BW flips the band
// as above.  The mean is floor((a + b) / 2) for integer samples and
// (a + b) / 2 in float32; the reference's numpy expression adds two uint8
// (or uint16) arrays first, which wraps above 255 (65535) -- that wrap is
// not reproduced.  DAT_SCL / DAT_OFFS / DAT_WTS are not applied (the
// reference does not apply them either).
#include "af_internal.h"

#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

using af::fail;

// fn(i) for i in [0, n), over up to 16 host threads (dynamic: a shared counter)
template <typename F>
void par_for(long long n, F fn) {
  if (n <= 0) return;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nt = (int)std::min<long long>(n, std::min(16u, hw));
  std::atomic<long long> next{0};
  auto body = [&]() {
    for (long long i; (i = next.fetch_add(1)) < n;) fn(i);
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(body);
  body();
  for (auto &t : th) t.join();
}

// positioned read of exactly n bytes (retrying short reads)
bool pread_all(int fd, void *dst, size_t n, off_t at) {
  uint8_t *p = (uint8_t *)dst;
  while (n) {
    const ssize_t k = ::pread(fd, p, n, at);
    if (k <= 0) return false;
    p += k;
    n -= (size_t)k;
    at += k;
  }
  return true;
}

// reverse the channel order of each row (es-byte elements)
void flip_rows(uint8_t *rows, size_t nrows, int nchans, size_t es) {
  const size_t blk = 4096;
  par_for((long long)((nrows + blk - 1) / blk), [&](long long b) {
    const size_t t1 = std::min(nrows, (size_t)(b + 1) * blk);
    for (size_t t = (size_t)b * blk; t < t1; ++t) {
      uint8_t *r = rows + t * nchans * es;
      for (int a = 0, c = nchans - 1; a < c; ++a, --c)
        for (size_t e = 0; e < es; ++e) std::swap(r[a * es + e], r[c * es + e]);
    }
  });
}

}  // namespace

struct af_fil_reader {
  int fd = -1;
  af_fil_info info{};
  size_t row_bytes = 0;
  bool flip = false;
};

struct af_psrfits_reader {
  int fd = -1;
  af_fil_info info{};
  long long data_off = 0;     // first byte of the SUBINT table's rows
  size_t naxis1 = 0;          // bytes per table row
  size_t nrows_tab = 0;       // table rows
  size_t col_off = 0;         // DATA column offset in a row
  int nsblk = 0, npol = 0;
  int es = 1;                 // bytes per stored sample
  int kind = 0;               // 0 uint8 ('B'), 1 uint16 ('I' + TZERO 32768), 2 int16 ('I'), 3 float32 ('E')
  bool flip = false;
};

namespace {

bool read_str(int fd, off_t &at, std::string &s) {
  int32_t n = 0;
  if (!pread_all(fd, &n, 4, at) || n < 0 || n > 4096) return false;
  at += 4;
  s.resize((size_t)n);
  if (n && !pread_all(fd, &s[0], (size_t)n, at)) return false;
  at += n;
  return true;
}

af_status parse_fil(int fd, af_fil_info &h, size_t &row_bytes, bool &flip) {
  off_t at = 0;
  std::string k;
  if (!read_str(fd, at, k) || k != "HEADER_START")
    return fail(AF_ERR_INVALID, "not a SIGPROC filterbank file (no HEADER_START)");
  int nchans = 0, nifs = 1, nbits = 8;
  double fch1 = 0, foff = 0, tstart = 0, tsamp = 0;
  auto rd_i = [&](int &v) {
    int32_t x = 0;
    const bool ok = pread_all(fd, &x, 4, at);
    at += 4;
    v = x;
    return ok;
  };
  auto rd_d = [&](double &v) {
    const bool ok = pread_all(fd, &v, 8, at);
    at += 8;
    return ok;
  };
  for (;;) {
    if (!read_str(fd, at, k)) return fail(AF_ERR_INVALID, "truncated filterbank header");
    if (k == "HEADER_END") break;
    bool ok = true;
    int iv;
    double dv;
    if (k == "nchans") ok = rd_i(nchans);
    else if (k == "nifs") ok = rd_i(nifs);
    else if (k == "nbits") ok = rd_i(nbits);
    else if (k == "fch1") ok = rd_d(fch1);
    else if (k == "foff") ok = rd_d(foff);
    else if (k == "tstart") ok = rd_d(tstart);
    else if (k == "tsamp") ok = rd_d(tsamp);
    else if (k == "source_name" || k == "rawdatafile") {
      std::string v;
      ok = read_str(fd, at, v);
    } else if (k == "telescope_id" || k == "machine_id" || k == "data_type" || k == "nbeams" ||
               k == "ibeam" || k == "barycentric" || k == "pulsarcentric" || k == "nsamples") {
      ok = rd_i(iv);
    } else if (k == "src_raj" || k == "src_dej" || k == "az_start" || k == "za_start" ||
               k == "refdm" || k == "period") {
      ok = rd_d(dv);
    } else {
      return fail(AF_ERR_INVALID, "unsupported filterbank header keyword '" + k + "'");
    }
    if (!ok) return fail(AF_ERR_INVALID, "truncated filterbank header");
  }
  if (nchans <= 0 || (nbits != 8 && nbits != 16 && nbits != 32))
    return fail(AF_ERR_INVALID, "unsupported filterbank layout (nchans/nbits)");
  if (nifs != 1) return fail(AF_ERR_INVALID, "only nifs == 1 is supported");
  const off_t end = ::lseek(fd, 0, SEEK_END);
  if (end < at) return fail(AF_ERR_INVALID, "truncated filterbank file");
  row_bytes = (size_t)nchans * (size_t)(nbits / 8);
  h = af_fil_info{};
  h.nchans = nchans;
  h.nbits = nbits;
  h.dtype = nbits == 8 ? AF_U8 : nbits == 16 ? AF_U16 : AF_U32;
  h.ndata = (uint64_t)((size_t)(end - at) / row_bytes);
  h.data_offset = (uint64_t)at;
  h.tstart = tstart;
  h.tsamp = tsamp;
  flip = foff < 0;
  h.fch1 = flip ? fch1 + (nchans - 1) * foff : fch1;
  h.foff = flip ? -foff : foff;
  h.flipped = flip ? 1 : 0;
  return AF_OK;
}

// ---- FITS -------------------------------------------------------------

struct Hdu {
  std::vector<std::pair<std::string, std::string>> cards;   // keyword, raw value
  long long data_off = 0;
  long long data_bytes = 0;
  const std::string *get(const std::string &k) const {
    for (const auto &c : cards)
      if (c.first == k) return &c.second;
    return nullptr;
  }
  bool num(const std::string &k, double &v) const {
    const std::string *s = get(k);
    if (!s) return false;
    std::string t = *s;
    for (char &ch : t)
      if (ch == 'D' || ch == 'd') ch = 'E';   // FORTRAN double exponents
    char *e = nullptr;
    v = std::strtod(t.c_str(), &e);
    return e != t.c_str();
  }
  std::string str(const std::string &k) const {
    const std::string *s = get(k);
    if (!s) return "";
    const size_t a = s->find('\''), b = s->rfind('\'');
    std::string v = (a != std::string::npos && b > a) ? s->substr(a + 1, b - a - 1) : *s;
    while (!v.empty() && v.back() == ' ') v.pop_back();
    return v;
  }
};

// Reads the header of the HDU at `at` (2880-byte blocks of 80-byte cards up
// to END) and the size of its data.
bool read_hdu(int fd, long long at, long long file_end, Hdu &h) {
  h.cards.clear();
  char block[2880];
  bool end = false;
  long long p = at;
  while (!end) {
    if (p + 2880 > file_end || !pread_all(fd, block, 2880, (off_t)p)) return false;
    p += 2880;
    for (int c = 0; c < 36 && !end; ++c) {
      const char *card = block + 80 * c;
      std::string key(card, 8);
      while (!key.empty() && key.back() == ' ') key.pop_back();
      if (key == "END") {
        end = true;
        break;
      }
      if (card[8] == '=' && card[9] == ' ') {
        std::string val(card + 10, 70);
        // the value ends at a '/' outside quotes
        bool q = false;
        size_t cut = val.size();
        for (size_t i = 0; i < val.size(); ++i) {
          if (val[i] == '\'') q = !q;
          else if (val[i] == '/' && !q) {
            cut = i;
            break;
          }
        }
        val = val.substr(0, cut);
        while (!val.empty() && val.front() == ' ') val.erase(val.begin());
        while (!val.empty() && val.back() == ' ') val.pop_back();
        h.cards.emplace_back(key, val);
      }
    }
  }
  h.data_off = p;
  double naxis = 0, bitpix = 8, pc = 0, gc = 1;
  h.num("NAXIS", naxis);
  h.num("BITPIX", bitpix);
  h.num("PCOUNT", pc);
  h.num("GCOUNT", gc);
  if (naxis < 0 || naxis > 999 || pc < 0 || gc < 0) return false;
  double n = naxis > 0 ? 1 : 0;
  for (int i = 1; i <= (int)naxis; ++i) {
    double v = 0;
    h.num("NAXIS" + std::to_string(i), v);
    if (v < 0) return false;
    n *= v;
  }
  const double bytes = std::fabs(bitpix) / 8 * gc * (n + pc);
  if (!(bytes >= 0) || bytes > 9.0e18) return false;
  h.data_bytes = (long long)bytes;
  return true;
}

// TFORM "rT" -> (repeat, type)
bool tform(const std::string &f, long long &rep, char &t) {
  size_t i = 0;
  while (i < f.size() && f[i] >= '0' && f[i] <= '9') ++i;
  rep = i ? std::atoll(f.substr(0, i).c_str()) : 1;
  if (i >= f.size()) return false;
  t = f[i];
  return true;
}

int tsize(char t) {
  switch (t) {
  case 'L': case 'B': case 'A': case 'X': return 1;   // X: bits, counted per byte below
  case 'I': return 2;
  case 'J': case 'E': return 4;
  case 'K': case 'D': case 'C': case 'P': return 8;
  case 'M': case 'Q': return 16;
  default: return -1;
  }
}

af_status parse_psrfits(int fd, af_psrfits_reader &r) {
  const long long file_end = (long long)::lseek(fd, 0, SEEK_END);
  Hdu h0, h;
  if (!read_hdu(fd, 0, file_end, h0) || h0.str("SIMPLE").empty())
    return fail(AF_ERR_INVALID, "not a FITS file");
  long long at = h0.data_off + (h0.data_bytes + 2879) / 2880 * 2880;
  bool found = false;
  while (at < file_end) {
    if (!read_hdu(fd, at, file_end, h)) break;
    if (h.str("XTENSION") == "BINTABLE" && h.str("EXTNAME") == "SUBINT") {
      found = true;
      break;
    }
    const long long next = h.data_off + (h.data_bytes + 2879) / 2880 * 2880;
    if (next <= at) break;
    at = next;
  }
  if (!found) return fail(AF_ERR_INVALID, "no SUBINT binary table (not PSRFITS search mode)");
  double nchan = 0, npol = 0, nsblk = 0, nbits = 0, tbin = 0, chan_bw = 0, naxis1 = 0, naxis2 = 0,
         tfields = 0;
  if (!h.num("NCHAN", nchan) || !h.num("NPOL", npol) || !h.num("NSBLK", nsblk) ||
      !h.num("NBITS", nbits) || !h.num("TBIN", tbin) || !h.num("CHAN_BW", chan_bw) ||
      !h.num("NAXIS1", naxis1) || !h.num("NAXIS2", naxis2) || !h.num("TFIELDS", tfields))
    return fail(AF_ERR_INVALID, "SUBINT header lacks NCHAN/NPOL/NSBLK/NBITS/TBIN/CHAN_BW");
  if (nchan < 1 || npol < 1 || nsblk < 1)
    return fail(AF_ERR_INVALID, "SUBINT with no channels, polarisations or samples");
  // the DATA column: its byte offset in a row and its element type
  long long off = 0;
  int col = 0;
  char type = 0;
  long long rep = 0;
  for (int i = 1; i <= (int)tfields; ++i) {
    long long rp;
    char t;
    if (!tform(h.str("TFORM" + std::to_string(i)), rp, t))
      return fail(AF_ERR_INVALID, "bad TFORM in SUBINT");
    const int sz = tsize(t);
    if (sz < 0) return fail(AF_ERR_INVALID, "unsupported TFORM type in SUBINT");
    const long long bytes = t == 'X' ? (rp + 7) / 8 : rp * sz;
    if (h.str("TTYPE" + std::to_string(i)) == "DATA") {
      col = i;
      type = t;
      rep = rp;
      break;
    }
    off += bytes;
  }
  if (!col) return fail(AF_ERR_INVALID, "SUBINT has no DATA column");
  double tzero = 0;
  h.num("TZERO" + std::to_string(col), tzero);
  if (type == 'B' && nbits == 8) r.kind = 0, r.es = 1;
  else if (type == 'I' && nbits == 16) r.kind = tzero == 32768.0 ? 1 : 2, r.es = 2;
  else if (type == 'E' && nbits == 32) r.kind = 3, r.es = 4;
  else
    return fail(AF_ERR_INVALID, "PSRFITS DATA of " + std::to_string((int)nbits) +
                                    "-bit samples is not supported (8-bit 'B', 16-bit 'I', "
                                    "32-bit float 'E')");
  if (r.kind == 2) return fail(AF_ERR_INVALID, "signed 16-bit PSRFITS samples are not supported");
  if (rep != (long long)nsblk * (long long)npol * (long long)nchan)
    return fail(AF_ERR_INVALID, "DATA column size is not NSBLK x NPOL x NCHAN");
  if (off + rep * r.es > (long long)naxis1)
    return fail(AF_ERR_INVALID, "DATA column extends past the SUBINT row");
  double obsfreq = 0, obsbw = 0, imjd = 0, smjd = 0, offs = 0;
  h0.num("OBSFREQ", obsfreq);
  h0.num("OBSBW", obsbw);
  h0.num("STT_IMJD", imjd);
  h0.num("STT_SMJD", smjd);
  h0.num("STT_OFFS", offs);
  r.data_off = h.data_off;
  r.naxis1 = (size_t)naxis1;
  r.nrows_tab = (size_t)naxis2;
  r.col_off = (size_t)off;
  r.nsblk = (int)nsblk;
  r.npol = (int)npol;
  if (r.data_off + (long long)(r.naxis1 * r.nrows_tab) > file_end)
    return fail(AF_ERR_INVALID, "truncated SUBINT table");
  af_fil_info &i = r.info;
  i = af_fil_info{};
  i.nchans = (int)nchan;
  i.nbits = (int)nbits;
  i.dtype = r.kind == 0 ? AF_U8 : r.kind == 1 ? AF_U16 : AF_F32;
  i.ndata = (uint64_t)r.nrows_tab * (uint64_t)r.nsblk;
  i.data_offset = (uint64_t)r.data_off;
  i.tsamp = tbin;
  i.tstart = imjd + smjd / 86400.0 + offs / 86400.0;
  double fch1 = obsfreq - obsbw / 2, foff = chan_bw;
  r.flip = foff < 0;
  if (r.flip) {
    foff = -foff;
    fch1 = fch1 - (nchan - 1) * foff;
  }
  i.fch1 = fch1;
  i.foff = foff;
  i.flipped = r.flip ? 1 : 0;
  i.npol = (int)npol;
  return AF_OK;
}

template <typename T>
inline T pol_mean(T a, T b) {
  return (T)(((uint32_t)a + (uint32_t)b) >> 1);
}
template <>
inline float pol_mean<float>(float a, float b) {
  return (a + b) / 2.0f;
}

inline uint16_t be16(const uint8_t *p) { return (uint16_t)((p[0] << 8) | p[1]); }
inline float bef32(const uint8_t *p) {
  const uint32_t u = ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// sample (pol, chan) of a block row: FITS big-endian; uint16 stored as
// int16 with TZERO 32768
template <typename T>
inline T sample(const uint8_t *blk, int kind);
template <>
inline uint8_t sample<uint8_t>(const uint8_t *p, int) { return *p; }
template <>
inline uint16_t sample<uint16_t>(const uint8_t *p, int) { return (uint16_t)(be16(p) ^ 0x8000u); }
template <>
inline float sample<float>(const uint8_t *p, int) { return bef32(p); }

template <typename T>
af_status psrfits_rows(const af_psrfits_reader &r, uint64_t row0, size_t nrows, T *dst) {
  const int nch = r.info.nchans, npol = r.npol, ns = r.nsblk, es = r.es;
  const size_t sub_bytes = (size_t)ns * npol * nch * es;
  const uint64_t s0 = row0 / (uint64_t)ns, s1 = (row0 + nrows + ns - 1) / (uint64_t)ns;
  std::atomic<bool> ok{true};
  par_for((long long)(s1 - s0), [&](long long j) {
    const uint64_t s = s0 + (uint64_t)j;
    std::vector<uint8_t> buf(sub_bytes);
    if (!pread_all(r.fd, buf.data(), sub_bytes,
                   (off_t)(r.data_off + (long long)s * (long long)r.naxis1 + (long long)r.col_off))) {
      ok = false;
      return;
    }
    for (int k = 0; k < ns; ++k) {
      const uint64_t t = s * ns + k;
      if (t < row0 || t >= row0 + nrows) continue;
      T *o = dst + (size_t)(t - row0) * nch;
      const uint8_t *p0 = buf.data() + (size_t)k * npol * nch * es;
      const uint8_t *p1 = p0 + (size_t)nch * es;
      for (int c = 0; c < nch; ++c) {
        const int oc = r.flip ? nch - 1 - c : c;
        const T a = sample<T>(p0 + (size_t)c * es, r.kind);
        o[oc] = npol >= 2 ? pol_mean<T>(a, sample<T>(p1 + (size_t)c * es, r.kind)) : a;
      }
    }
  });
  return ok ? AF_OK : fail(AF_ERR_RUNTIME, "PSRFITS read failed");
}

}  // namespace

extern "C" {

af_status af_fil_open(const char *path, af_fil_reader **out, af_fil_info *info) {
  if (!path || !out) return fail(AF_ERR_INVALID, "null argument");
  *out = nullptr;
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) return fail(AF_ERR_RUNTIME, std::string("Can not open file ") + path);
  af_fil_reader *r = new af_fil_reader;
  r->fd = fd;
  af_status s = parse_fil(fd, r->info, r->row_bytes, r->flip);
  if (s != AF_OK) {
    ::close(fd);
    delete r;
    return s;
  }
  if (info) *info = r->info;
  *out = r;
  return AF_OK;
}

af_status af_fil_read(af_fil_reader *r, uint64_t row0, size_t nrows, void *dst) {
  if (!r || (!dst && nrows)) return fail(AF_ERR_INVALID, "null argument");
  if (row0 > r->info.ndata || nrows > r->info.ndata - row0)
    return fail(AF_ERR_INVALID, "rows beyond the end of the file");
  uint8_t *d = (uint8_t *)dst;
  const size_t rb = r->row_bytes;
  // blocks of rows over the threads, one positioned read each
  const size_t blk = std::max<size_t>(1, ((size_t)8 << 20) / rb);
  const long long nblk = (long long)((nrows + blk - 1) / blk);
  std::atomic<bool> ok{true};
  par_for(nblk, [&](long long b) {
    const size_t a = (size_t)b * blk, n = std::min(blk, nrows - a);
    if (!pread_all(r->fd, d + a * rb, n * rb,
                   (off_t)(r->info.data_offset + (row0 + a) * (uint64_t)rb)))
      ok = false;
  });
  if (!ok) return fail(AF_ERR_RUNTIME, "filterbank read failed");
  if (r->flip) flip_rows(d, nrows, r->info.nchans, (size_t)(r->info.nbits / 8));
  return AF_OK;
}

af_status af_fil_close(af_fil_reader *r) {
  if (!r) return AF_OK;
  if (r->fd >= 0) ::close(r->fd);
  delete r;
  return AF_OK;
}

af_status af_psrfits_open(const char *path, af_psrfits_reader **out, af_fil_info *info) {
  if (!path || !out) return fail(AF_ERR_INVALID, "null argument");
  *out = nullptr;
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) return fail(AF_ERR_RUNTIME, std::string("Can not open file ") + path);
  af_psrfits_reader *r = new af_psrfits_reader;
  r->fd = fd;
  af_status s = parse_psrfits(fd, *r);
  if (s != AF_OK) {
    ::close(fd);
    delete r;
    return s;
  }
  if (info) *info = r->info;
  *out = r;
  return AF_OK;
}

af_status af_psrfits_read(af_psrfits_reader *r, uint64_t row0, size_t nrows, void *dst) {
  if (!r || (!dst && nrows)) return fail(AF_ERR_INVALID, "null argument");
  if (row0 > r->info.ndata || nrows > r->info.ndata - row0)
    return fail(AF_ERR_INVALID, "rows beyond the end of the file");
  if (!nrows) return AF_OK;
  switch (r->kind) {
  case 0: return psrfits_rows<uint8_t>(*r, row0, nrows, (uint8_t *)dst);
  case 1: return psrfits_rows<uint16_t>(*r, row0, nrows, (uint16_t *)dst);
  default: return psrfits_rows<float>(*r, row0, nrows, (float *)dst);
  }
}

af_status af_psrfits_close(af_psrfits_reader *r) {
  if (!r) return AF_OK;
  if (r->fd >= 0) ::close(r->fd);
  delete r;
  return AF_OK;
}

}  // extern "C"
