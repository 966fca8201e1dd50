// survey.cpp -- the survey description and the data files of a run: model and
// frequencies (emf_init), acquisition tables (acq_init), and the ASCII data
// files emf_/obs_/syn_<shot>.txt and rmse_<shot>.txt.  Parameter names,
// defaults, checks, messages and file formats are those of
// src/emf_init_free.c, src/acq_init_free.c and src/read_write_data.c; the
// text formats have to be byte-identical so that the files of either program
// can be read by the other.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "arrays.h"
#include "emmi_cli.h"

using namespace emmi_cli;

namespace {
constexpr double kPi = 3.14159265358979323846264;
char g_default_channel[] = "Ex";

// one model file: nx*ny*nz float32 (src/emf_init_free.c:87-107)
void read_model(const char* key, const char* name, float*** a, int nxyz) {
  FILE* fp = fopen(name, "rb");
  if (!fp) emmi_cli_err("cannot open file %s=%s\n", key, name);
  const int got = static_cast<int>(fread(&a[0][0][0], sizeof(float), nxyz, fp));
  fclose(fp);
  if (got != nxyz) emmi_cli_err("size not match %s: file=%d nx*ny*nz=%d\n", key, got, nxyz);
}

// channels= list, "Ex" when absent (src/emf_init_free.c:27-46)
int read_channels(const char* key, char*** out) {
  int n = countparval(const_cast<char*>(key));
  if (n) {
    *out = alloc1<char*>(n);
    getparstringarray(const_cast<char*>(key), *out);
  } else {
    n = 1;
    *out = alloc1<char*>(1);
    (*out)[0] = g_default_channel;
  }
  return n;
}

struct Station {
  float x, y, z, hd, pit;
};

// a source or receiver table: one header line, then "x y z azimuth dip index"
// rows until end of file (src/acq_init_free.c:83-109, 164-190); the running
// bounds are widened by every row
std::vector<Station> read_stations(const char* name, const char* what, float* lo, float* hi) {
  FILE* fp = fopen(name, "r");
  if (!fp) emmi_cli_err("file %s= missing", what);
  if (fscanf(fp, "%*[^\n]\n") == EOF) {
  }
  std::vector<Station> v;
  while (true) {
    Station s;
    int idx;
    const int got = fscanf(fp, "%f %f %f %f %f %d", &s.x, &s.y, &s.z, &s.hd, &s.pit, &idx);
    if (got == EOF) break;
    if (got != 6) emmi_cli_err("%s=%s: row %zu is not 'x y z azimuth dip index'", what, name, v.size() + 1);
    v.push_back(s);
    lo[0] = std::min(lo[0], s.x);
    hi[0] = std::max(hi[0], s.x);
    lo[1] = std::min(lo[1], s.y);
    hi[1] = std::max(hi[1], s.y);
    lo[2] = std::min(lo[2], s.z);
    hi[2] = std::max(hi[2], s.z);
  }
  fclose(fp);
  return v;
}

void check_bounds(const acq_t* acq, const float* lo, const float* hi, const char* what) {
  if (lo[0] < acq->x1min) emmi_cli_err("%s location: x<x1min", what);
  if (lo[1] < acq->x2min) emmi_cli_err("%s location: y<x2min", what);
  if (lo[2] < acq->x3min) emmi_cli_err("%s location: z<x3min", what);
  if (hi[0] > acq->x1max) emmi_cli_err("%s location: x>x1max", what);
  if (hi[1] > acq->x2max) emmi_cli_err("%s location: y>x2max", what);
  if (hi[2] > acq->x3max) emmi_cli_err("%s location: z>x3max", what);
}

float deg2rad(float d) { return static_cast<float>(kPi * d / 180.); }
}  // namespace

extern "C" {

void emf_init(emf_t* emf) {
  if (iproc == 0) printf("---------- emf_init -----------\n");
  emf->nfreq = countparval(const_cast<char*>("freqs"));
  if (!emf->nfreq) emmi_cli_err("Need freqs= vector");
  emf->freqs = alloc1<float>(emf->nfreq);
  emf->omegas = alloc1<float>(emf->nfreq);
  getparfloat(const_cast<char*>("freqs"), emf->freqs);
  for (int i = 0; i < emf->nfreq; ++i) {
    emf->omegas[i] = static_cast<float>(2. * kPi * emf->freqs[i]);
    if (emf->verb) printf("freq[%d]=%g \n", i, emf->freqs[i]);
  }
  emf->nchsrc = read_channels("chsrc", &emf->chsrc);
  emf->nchrec = read_channels("chrec", &emf->chrec);
  if (emf->verb) {
    printf("Active source channels:");
    for (int i = 0; i < emf->nchsrc; ++i) printf(" %s", emf->chsrc[i]);
    printf("\nActive recever channels:");
    for (int i = 0; i < emf->nchrec; ++i) printf(" %s", emf->chrec[i]);
    printf("\n");
  }
  if (!getparint(const_cast<char*>("reciprocity"), &emf->reciprocity)) emf->reciprocity = 0;
  if (!getparint(const_cast<char*>("nx"), &emf->nx)) emmi_cli_err("Need nx= ");
  if (!getparint(const_cast<char*>("ny"), &emf->ny)) emmi_cli_err("Need ny= ");
  if (!getparint(const_cast<char*>("nz"), &emf->nz)) emmi_cli_err("Need nz= ");
  if (!getparfloat(const_cast<char*>("dx"), &emf->dx)) emmi_cli_err("Need dx= ");
  if (!getparfloat(const_cast<char*>("dy"), &emf->dy)) emmi_cli_err("Need dy= ");
  if (!getparfloat(const_cast<char*>("dz"), &emf->dz)) emmi_cli_err("Need dz= ");
  if (!getparfloat(const_cast<char*>("ox"), &emf->ox)) emmi_cli_err("Need ox= ");
  if (!getparfloat(const_cast<char*>("oy"), &emf->oy)) emmi_cli_err("Need oy= ");
  if (!getparfloat(const_cast<char*>("oz"), &emf->oz)) emf->oz = 0;
  if (!getparfloat(const_cast<char*>("d1min"), &emf->d1min)) emf->d1min = 100;
  if (!getparfloat(const_cast<char*>("d2min"), &emf->d2min)) emf->d2min = 100;
  if (!getparfloat(const_cast<char*>("d3min"), &emf->d3min)) emf->d3min = 20;
  if (!getparint(const_cast<char*>("n1"), &emf->n1)) emf->n1 = 128;
  if (!getparint(const_cast<char*>("n2"), &emf->n2)) emf->n2 = 128;
  if (!getparint(const_cast<char*>("n3"), &emf->n3)) emf->n3 = 128;
  if (emf->nx <= 0 || emf->ny <= 0 || emf->nz <= 0) emmi_cli_err("nx, ny, nz must be positive");
  if (emf->verb) {
    printf("[nx, ny, nz]=[%d, %d, %d]\n", emf->nx, emf->ny, emf->nz);
    printf("[dx, dy, dz]=[%g, %g, %g]\n", emf->dx, emf->dy, emf->dz);
    printf("[ox, oy, oz]=[%g, %g, %g]\n", emf->ox, emf->oy, emf->oz);
    printf("[d1min, d2min, d3min]=[%g, %g, %g]\n", emf->d1min, emf->d2min, emf->d3min);
    printf("GMG grid size: [n1, n2, n3]=[%d, %d, %d]\n", emf->n1, emf->n2, emf->n3);
  }
  char *f11, *f22, *f33;
  if (!getparstring(const_cast<char*>("frho11"), &f11)) emmi_cli_err("Need frho11= ");
  if (!getparstring(const_cast<char*>("frho22"), &f22)) emmi_cli_err("Need frho22= ");
  if (!getparstring(const_cast<char*>("frho33"), &f33)) emmi_cli_err("Need frho33= ");
  const int nxyz = emf->nx * emf->ny * emf->nz;
  emf->rho11 = alloc3<float>(emf->nx, emf->ny, emf->nz);
  emf->rho22 = alloc3<float>(emf->nx, emf->ny, emf->nz);
  emf->rho33 = alloc3<float>(emf->nx, emf->ny, emf->nz);
  read_model("frho11", f11, emf->rho11, nxyz);
  read_model("frho22", f22, emf->rho22, nxyz);
  read_model("frho33", f33, emf->rho33, nxyz);
  if (!getparfloat(const_cast<char*>("lextend"), &emf->lextend)) emf->lextend = 50e3;
  if (!getparfloat(const_cast<char*>("rho_air"), &emf->rho_air)) emf->rho_air = 1e8;
  if (!getparint(const_cast<char*>("nb"), &emf->nb)) emf->nb = 10;
  // rho_water = the smallest resistivity of the three inputs
  // (src/emf_init_free.c:115-128)
  const float *r11 = &emf->rho11[0][0][0], *r22 = &emf->rho22[0][0][0], *r33 = &emf->rho33[0][0][0];
  float vmax = r11[0], vmin = r11[0];
  for (int i = 0; i < nxyz; ++i) {
    vmax = std::max(vmax, r11[i]);
    vmin = std::min(vmin, r11[i]);
    vmax = std::max(vmax, r22[i]);
    vmin = std::min(vmin, r22[i]);
    vmax = std::max(vmax, r33[i]);
    vmin = std::min(vmin, r33[i]);
  }
  emf->rho_water = vmin;
  if (emf->verb) {
    printf("input model [rhomin, rhomax]=[%g, %g]\n", vmin, vmax);
    printf("rho_air=%g\n", emf->rho_air);
    printf("rho_water=%g\n", emf->rho_water);
    printf("extended distance on each side: lextend=%g\n", emf->lextend);
  }
}

void emf_free(emf_t* emf) {
  free(emf->chsrc);
  free(emf->chrec);
  emf->chsrc = emf->chrec = nullptr;
  free(emf->freqs);
  free(emf->omegas);
  emf->freqs = emf->omegas = nullptr;
  free3(emf->rho11);
  free3(emf->rho22);
  free3(emf->rho33);
  emf->rho11 = emf->rho22 = emf->rho33 = nullptr;
}

void acq_init(acq_t* acq, emf_t* emf) {
  if (iproc == 0) printf("---------acq_init ---------\n");
  char *fsrc, *frec, *fsrcrec;
  if (!getparstring(const_cast<char*>("fsrc"), &fsrc)) emmi_cli_err("Need fsrc= ");
  if (!getparstring(const_cast<char*>("frec"), &frec)) emmi_cli_err("Need frec= ");
  if (!getparstring(const_cast<char*>("fsrcrec"), &fsrcrec)) emmi_cli_err("Need fsrcrec= ");
  if (!getparfloat(const_cast<char*>("x1min"), &acq->x1min)) acq->x1min = emf->ox;
  if (!getparfloat(const_cast<char*>("x1max"), &acq->x1max)) acq->x1max = emf->ox + emf->nx * emf->dx;
  if (!getparfloat(const_cast<char*>("x2min"), &acq->x2min)) acq->x2min = emf->oy;
  if (!getparfloat(const_cast<char*>("x2max"), &acq->x2max)) acq->x2max = emf->oy + emf->ny * emf->dy;
  if (!getparfloat(const_cast<char*>("x3min"), &acq->x3min)) acq->x3min = emf->oz;
  if (!getparfloat(const_cast<char*>("x3max"), &acq->x3max)) acq->x3max = emf->oz + emf->nz * emf->dz;
  if (emf->ox > acq->x1min || emf->ox + emf->dx * emf->nx < acq->x1max)
    emmi_cli_err("x - sources/receivers from acquisition file are out of domain");
  if (emf->oy > acq->x2min || emf->oy + emf->dy * emf->ny < acq->x2max)
    emmi_cli_err("y - sources/receivers from acquisition file are out of domain");
  if (emf->oz > acq->x3min || emf->oz + emf->dz * emf->nz < acq->x3max)
    emmi_cli_err("z - sources/receivers from acquisition file are out of domain");
  if (!getparint(const_cast<char*>("nsubsrc"), &acq->nsubsrc)) acq->nsubsrc = 1;
  if (!getparint(const_cast<char*>("nsubrec"), &acq->nsubrec)) acq->nsubrec = 1;
  if (!getparfloat(const_cast<char*>("lensrc"), &acq->lensrc)) acq->lensrc = 275.;
  if (!getparfloat(const_cast<char*>("lenrec"), &acq->lenrec)) acq->lenrec = 8.;
  if (std::fabs(acq->x1max - acq->x1min - emf->dx * emf->nx) > 1e-15)
    emmi_cli_err("inconsistent input: x1max-x1min!=dx*nx");
  if (std::fabs(acq->x2max - acq->x2min - emf->dy * emf->ny) > 1e-15)
    emmi_cli_err("inconsistent input: x2max-x2min!=dy*ny");

  // step 1: every source position; bounds over sources
  float lo[3] = {acq->x1min, acq->x2min, acq->x3min}, hi[3] = {acq->x1max, acq->x2max, acq->x3max};
  const std::vector<Station> src = read_stations(fsrc, "fsrc", lo, hi);
  acq->nsrc_total = static_cast<int>(src.size());
  check_bounds(acq, lo, hi, "source");

  // the shot of each rank: shots= (one per rank) or 1..nproc
  // (src/acq_init_free.c:118-141)
  acq->shot_idx = alloc1<int>(nproc);
  const int nshots = countparval(const_cast<char*>("shots"));
  if (nshots > 0) {
    if (nshots != nproc) emmi_cli_err("shots length (%d) must equal nproc (%d)", nshots, nproc);
    getparint(const_cast<char*>("shots"), acq->shot_idx);
    for (int i = 0; i < nproc; ++i)
      if (acq->shot_idx[i] < 1 || acq->shot_idx[i] > acq->nsrc_total)
        emmi_cli_err("shots[%d]=%d out of range [1,%d]", i, acq->shot_idx[i], acq->nsrc_total);
  } else {
    if (nproc > acq->nsrc_total)
      emmi_cli_err("nproc=%d > nsrc_total=%d; reduce nproc or provide shots=", nproc, acq->nsrc_total);
    for (int i = 0; i < nproc; ++i) acq->shot_idx[i] = i + 1;
  }
  const int idx = acq->shot_idx[iproc];

  // the receivers of this shot from the source-receiver table; with
  // reciprocity the shot is a receiver gather (src/acq_init_free.c:148-170)
  std::vector<int> rec_idx;
  {
    FILE* fp = fopen(fsrcrec, "r");
    if (!fp) emmi_cli_err("file fsrcrec= missing!");
    if (fscanf(fp, "%*[^\n]\n") == EOF) {
    }
    int isrc, irec, got;
    while ((got = fscanf(fp, "%d %d", &isrc, &irec)) != EOF) {
      if (got != 2) emmi_cli_err("fsrcrec=%s: a row is not 'isrc irec'", fsrcrec);
      if (emf->reciprocity) {
        if (irec == idx) rec_idx.push_back(isrc);
      } else if (isrc == idx) {
        rec_idx.push_back(irec);
      }
    }
    fclose(fp);
  }
  acq->nrec = static_cast<int>(rec_idx.size());

  // step 2: every receiver position; bounds continue over receivers
  const std::vector<Station> rec = read_stations(frec, "frec", lo, hi);
  acq->nrec_total = static_cast<int>(rec.size());
  check_bounds(acq, lo, hi, "receiver");

  const std::vector<Station>& as_src = emf->reciprocity ? rec : src;
  const std::vector<Station>& as_rec = emf->reciprocity ? src : rec;
  acq->nsrc = 1;
  acq->src_x1 = alloc1<float>(1);
  acq->src_x2 = alloc1<float>(1);
  acq->src_x3 = alloc1<float>(1);
  acq->src_azimuth = alloc1<float>(1);
  acq->src_dip = alloc1<float>(1);
  {
    const int i = idx - 1;
    if (i < 0 || i >= static_cast<int>(as_src.size())) emmi_cli_err("shot %d has no position in the tables", idx);
    const Station& s = as_src[i];
    acq->src_x1[0] = s.x;
    acq->src_x2[0] = s.y;
    acq->src_x3[0] = s.z;
    acq->src_azimuth[0] = deg2rad(s.hd);
    acq->src_dip[0] = deg2rad(s.pit);
  }
  acq->rec_x1 = alloc1<float>(acq->nrec);
  acq->rec_x2 = alloc1<float>(acq->nrec);
  acq->rec_x3 = alloc1<float>(acq->nrec);
  acq->rec_azimuth = alloc1<float>(acq->nrec);
  acq->rec_dip = alloc1<float>(acq->nrec);
  for (int r = 0; r < acq->nrec; ++r) {
    const int i = rec_idx[r] - 1;
    if (i < 0 || i >= static_cast<int>(as_rec.size()))
      emmi_cli_err("table entry %d of shot %d has no position", rec_idx[r], idx);
    const Station& s = as_rec[i];
    acq->rec_x1[r] = s.x;
    acq->rec_x2[r] = s.y;
    acq->rec_x3[r] = s.z;
    acq->rec_azimuth[r] = deg2rad(s.hd);
    acq->rec_dip[r] = deg2rad(s.pit);
  }
  printf("****** isrc=%d nrec=%d, [x,y,z,azimuth,dip]=[%g, %g, %g, %g, %g]\n", idx, acq->nrec, acq->src_x1[0],
         acq->src_x2[0], acq->src_x3[0], 180 * acq->src_azimuth[0] / kPi, 180. * acq->src_dip[0] / kPi);
  if (emf->verb) {
    printf("lensrc=%g, nsubsrc=%d\n", acq->lensrc, acq->nsubsrc);
    printf("lenrec=%g, nsubrec=%d\n", acq->lenrec, acq->nsubrec);
    printf("original domain [x1min, x1max]=[%g, %g]\n", acq->x1min, acq->x1max);
    printf("original domain [x2min, x2max]=[%g, %g]\n", acq->x2min, acq->x2max);
    printf("original domain [x3min, x3max]=[%g, %g]\n", acq->x3min, acq->x3max);
    printf("nsrc_total=%d\n", acq->nsrc_total);
    printf("nrec_total=%d\n", acq->nrec_total);
  }
}

void acq_free(acq_t* acq) {
  float** arrs[] = {&acq->src_x1, &acq->src_x2, &acq->src_x3, &acq->src_azimuth, &acq->src_dip,
                    &acq->rec_x1, &acq->rec_x2, &acq->rec_x3, &acq->rec_azimuth, &acq->rec_dip};
  for (float** a : arrs) {
    free(*a);
    *a = nullptr;
  }
  free(acq->shot_idx);
  acq->shot_idx = nullptr;
}

void write_data(acq_t* acq, emf_t* emf, char* fname, emmi_c64*** d) {
  FILE* fp = fopen(fname, "w");
  if (!fp) emmi_cli_err("error opening file for writing");
  fprintf(fp, "iTx \tiRx    chrec  frequency/Hz    Real{E/H}     Imag{E/H}\n");
  const int isrc = acq->shot_idx[iproc];
  for (int c = 0; c < emf->nchrec; ++c)
    for (int f = 0; f < emf->nfreq; ++f)
      for (int r = 0; r < acq->nrec; ++r)
        fprintf(fp, "%d \t %d \t %s \t %g \t %e \t %e\n", isrc, r + 1, emf->chrec[c], emf->freqs[f], d[c][f][r].re,
                d[c][f][r].im);
  fclose(fp);
}

void read_data(acq_t* acq, emf_t* emf) {
  char fname[32];
  snprintf(fname, sizeof fname, "emf_%04d.txt", acq->shot_idx[iproc]);
  FILE* fp = fopen(fname, "r");
  if (!fp) emmi_cli_err("error opening file for reading");
  if (fscanf(fp, "%*[^\n]\n") == EOF) {
  }
  int isrc, irec;
  float re, im, freq;
  char ch[64];
  int got;
  while ((got = fscanf(fp, "%d \t %d \t %63s \t %g \t %e \t %e\n", &isrc, &irec, ch, &freq, &re, &im)) != EOF) {
    if (got != 6) emmi_cli_err("%s: a line is not 'iTx iRx chrec frequency real imag'", fname);
    if (acq->shot_idx[iproc] != isrc || irec > acq->nrec || irec < 1) continue;
    for (int c = 0; c < emf->nchrec; ++c) {
      if (strcmp(emf->chrec[c], ch) != 0) continue;
      for (int f = 0; f < emf->nfreq; ++f)
        if (std::fabs(emf->freqs[f] - freq) < 1e-4) {
          emf->dobs_fd[c][f][irec - 1] = {re, im};
          break;
        }
      break;
    }
  }
  fclose(fp);
}

void write_rmse(acq_t* acq, emf_t* emf) {
  char fname[32];
  snprintf(fname, sizeof fname, "rmse_%04d.txt", acq->shot_idx[iproc]);
  FILE* fp = fopen(fname, "w");
  if (!fp) emmi_cli_err("error opening file for writing");
  fprintf(fp, "iTx iRx chrec frequency/Hz RMSE x y z\n");
  const int isrc = acq->shot_idx[iproc];
  for (int c = 0; c < emf->nchrec; ++c)
    for (int f = 0; f < emf->nfreq; ++f)
      for (int r = 0; r < acq->nrec; ++r)
        fprintf(fp, "%d %d %s %g %e %g %g %g\n", isrc, r + 1, emf->chrec[c], emf->freqs[f], emf->rmse[c][f][r],
                acq->rec_x1[r], acq->rec_x2[r], acq->rec_x3[r]);
  fclose(fp);
}
}
