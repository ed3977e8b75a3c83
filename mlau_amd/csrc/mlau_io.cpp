// mlau_io.cpp -- the reference's output and backup file layouts (host code).
// Text formats and binary record orders follow the drivers cited in
// include/mlau_io.h; file names are built with the same zero-padded counters.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mlau_io.h"

namespace mlau {
int fail(int code, const std::string &msg);
}
using mlau::fail;

namespace {

std::string path(const char *dir, const char *name) { return std::string(dir) + "/" + name; }

std::string numbered(const char *dir, const char *stem, int a, int b)
{
  char buf[64];
  if (b < 0) snprintf(buf, sizeof buf, "%s_%05d.dat", stem, a);
  else snprintf(buf, sizeof buf, "%s_%05d_%05d.dat", stem, a, b);
  return path(dir, buf);
}

long long cells(const mlau_grid *g)
{
  long long n = g->nx;
  if (g->dim >= 2) n *= g->ny;
  if (g->dim == 3) n *= g->nz;
  return n;
}

int nfields(const mlau_grid *g) { return g->dim == 1 ? 7 : 8; }

int check(const char *dir, const mlau_grid *g)
{
  if (!dir || !g) return fail(MLAU_ERR_ARG, "null argument");
  if (g->dim < 1 || g->dim > 3 || g->nx < 1 || (g->dim >= 2 && g->ny < 1) || (g->dim == 3 && g->nz < 1))
    return fail(MLAU_ERR_ARG, "bad grid");
  return MLAU_OK;
}

bool write_text(const std::string &p, const char *mode, const std::string &txt)
{
  FILE *f = fopen(p.c_str(), mode);
  if (!f) return false;
  const bool ok = fwrite(txt.data(), 1, txt.size(), f) == txt.size();
  return fclose(f) == 0 && ok;
}

std::string fmt(const char *f, double a)
{
  char buf[64];
  snprintf(buf, sizeof buf, f, a);
  return buf;
}

bool write_coords(const std::string &p, const double *v, int n)
{
  std::string s;
  for (int i = 0; i < n; i++) s += fmt("%.12f\n", v[i]);
  return write_text(p, "w", s);
}

}  // namespace

extern "C" {

int mlau_io_write_grid(const char *dir, const mlau_grid *g, const double *x, const double *y, const double *z,
                       int rank, int ix, int iy, int iz)
{
  int rc = check(dir, g);
  if (rc) return rc;
  const double *v[3] = {x, y, z};
  const int n[3] = {g->nx, g->ny, g->nz};
  const int idx[3] = {ix, iy, iz};
  const char *stem[3] = {"x", "y", "z"};
  for (int a = 0; a < g->dim; a++) {
    if (!v[a]) return fail(MLAU_ERR_ARG, "missing coordinate array");
    const std::string p = rank < 0 ? path(dir, (std::string(stem[a]) + ".dat").c_str()) : numbered(dir, stem[a], idx[a], -1);
    if (!write_coords(p, v[a], n[a])) return fail(MLAU_ERR_STATE, "cannot write " + p);
  }
  return MLAU_OK;
}

int mlau_io_write_output(const char *dir, const mlau_grid *g, int first, int cnt, double tim,
                         const double *const *fields, int fmtc, int rank)
{
  int rc = check(dir, g);
  if (rc) return rc;
  if (!fields || (fmtc != MLAU_IO_F64 && fmtc != MLAU_IO_F32)) return fail(MLAU_ERR_ARG, "bad fields or format");
  if (rank <= 0) {
    bool ok = true;
    const char *tf = g->dim == 1 ? "%.12f\n" : "%f\n";
    if (first) {
      if (g->dim == 1) {
        char buf[96];
        snprintf(buf, sizeof buf, "%.12f %.12f\n", g->bx, g->gamma);
        ok = write_text(path(dir, "params.dat"), "w", buf) &&
             write_text(path(dir, "xoff.dat"), "w", std::to_string(g->xoff) + "\n");
      } else {
        std::string off = std::to_string(g->xoff) + " " + std::to_string(g->yoff);
        if (g->dim == 3) off += " " + std::to_string(g->zoff);
        ok = write_text(path(dir, "params.dat"), "w", fmt("%f\n", g->gamma)) &&
             write_text(path(dir, "offsets.dat"), "w", off + "\n");
      }
      ok = ok && write_text(path(dir, "t.dat"), "w", fmt(tf, tim));
    } else {
      ok = write_text(path(dir, "t.dat"), "a", fmt(tf, tim));
    }
    if (!ok) return fail(MLAU_ERR_STATE, std::string("cannot write the text files in ") + dir);
  }
  const std::string p = numbered(dir, "outdat", cnt, rank < 0 ? -1 : rank);
  FILE *f = fopen(p.c_str(), "wb");
  if (!f) return fail(MLAU_ERR_STATE, "cannot open " + p);
  const long long nall = cells(g);
  bool ok = true;
  std::vector<float> buf(fmtc == MLAU_IO_F32 ? nall : 0);
  for (int k = 0; k < nfields(g) && ok; k++) {
    if (!fields[k]) {
      ok = false;
      break;
    }
    if (fmtc == MLAU_IO_F64) {
      ok = fwrite(fields[k], sizeof(double), nall, f) == (size_t)nall;
    } else {
      for (long long i = 0; i < nall; i++) buf[i] = (float)fields[k][i];
      ok = fwrite(buf.data(), sizeof(float), nall, f) == (size_t)nall;
    }
  }
  if (fclose(f) != 0) ok = false;
  return ok ? MLAU_OK : fail(MLAU_ERR_STATE, "cannot write " + p);
}

int mlau_io_read_output(const char *dir, const mlau_grid *g, int cnt, int fmtc, int rank, double *const *fields)
{
  int rc = check(dir, g);
  if (rc) return rc;
  if (!fields || (fmtc != MLAU_IO_F64 && fmtc != MLAU_IO_F32)) return fail(MLAU_ERR_ARG, "bad fields or format");
  const std::string p = numbered(dir, "outdat", cnt, rank < 0 ? -1 : rank);
  FILE *f = fopen(p.c_str(), "rb");
  if (!f) return fail(MLAU_ERR_STATE, "cannot open " + p);
  const long long nall = cells(g);
  bool ok = true;
  std::vector<float> buf(fmtc == MLAU_IO_F32 ? nall : 0);
  for (int k = 0; k < nfields(g) && ok; k++) {
    if (!fields[k]) {
      ok = false;
      break;
    }
    if (fmtc == MLAU_IO_F64) {
      ok = fread(fields[k], sizeof(double), nall, f) == (size_t)nall;
    } else {
      ok = fread(buf.data(), sizeof(float), nall, f) == (size_t)nall;
      for (long long i = 0; ok && i < nall; i++) fields[k][i] = buf[i];
    }
  }
  fclose(f);
  return ok ? MLAU_OK : fail(MLAU_ERR_STATE, "short or unreadable record " + p);
}

// argument order ro,mx,my,mz,en,bx,by,bz -> backup order ro,mx,my,mz,bx,by,bz,en
static const int kBkupOrder[8] = {0, 1, 2, 3, 5, 6, 7, 4};

int mlau_io_bkup_save(const char *dir, long long nall, int n, int cnt, double tim, double dt, double trec, int rank,
                      const double *const *fields)
{
  if (!dir || !fields || nall < 1) return fail(MLAU_ERR_ARG, "bad argument");
  const int r = rank < 0 ? 0 : rank;
  if (r == 0) {
    char buf[64];
    snprintf(buf, sizeof buf, "%d %d\n", n, cnt);
    if (!write_text(path(dir, "bkup_stamp.dat"), "w", buf)) return fail(MLAU_ERR_STATE, "cannot write bkup_stamp.dat");
    const double v[3] = {tim, dt, trec};
    FILE *f = fopen(path(dir, "bkup_time.dat").c_str(), "wb");
    if (!f) return fail(MLAU_ERR_STATE, "cannot write bkup_time.dat");
    const bool ok = fwrite(v, sizeof(double), 3, f) == 3;
    if (fclose(f) != 0 || !ok) return fail(MLAU_ERR_STATE, "cannot write bkup_time.dat");
  }
  const std::string p = numbered(dir, "bkup_data", r, -1);
  FILE *f = fopen(p.c_str(), "wb");
  if (!f) return fail(MLAU_ERR_STATE, "cannot open " + p);
  bool ok = true;
  for (int k = 0; k < 8 && ok; k++) {
    const double *src = fields[kBkupOrder[k]];
    ok = src && fwrite(src, sizeof(double), nall, f) == (size_t)nall;
  }
  if (fclose(f) != 0) ok = false;
  return ok ? MLAU_OK : fail(MLAU_ERR_STATE, "cannot write " + p);
}

int mlau_io_bkup_load(const char *dir, long long nall, int rank, int *n, int *cnt, double *tim, double *dt,
                      double *trec, double *const *fields)
{
  if (!dir || !fields || nall < 1 || !n || !cnt || !tim || !dt || !trec) return fail(MLAU_ERR_ARG, "bad argument");
  FILE *f = fopen(path(dir, "bkup_stamp.dat").c_str(), "r");
  if (!f) return MLAU_NO_BACKUP;
  int v[2] = {0, 0};
  const int got = fscanf(f, "%d %d", &v[0], &v[1]);
  fclose(f);
  if (got != 2) return fail(MLAU_ERR_STATE, "malformed bkup_stamp.dat");
  double t[3];
  f = fopen(path(dir, "bkup_time.dat").c_str(), "rb");
  if (!f) return fail(MLAU_ERR_STATE, "missing bkup_time.dat");
  const bool okt = fread(t, sizeof(double), 3, f) == 3;
  fclose(f);
  if (!okt) return fail(MLAU_ERR_STATE, "short bkup_time.dat");
  const std::string p = numbered(dir, "bkup_data", rank < 0 ? 0 : rank, -1);
  f = fopen(p.c_str(), "rb");
  if (!f) return fail(MLAU_ERR_STATE, "missing " + p);
  bool ok = true;
  for (int k = 0; k < 8 && ok; k++) {
    double *dst = fields[kBkupOrder[k]];
    ok = dst && fread(dst, sizeof(double), nall, f) == (size_t)nall;
  }
  fclose(f);
  if (!ok) return fail(MLAU_ERR_STATE, "short or unreadable " + p);
  *n = v[0];
  *cnt = v[1];
  *tim = t[0];
  *dt = t[1];
  *trec = t[2];
  return MLAU_OK;
}

}  // extern "C"
