// CPU oracle for the trax track-seeding chain (TEST INFRASTRUCTURE ONLY).
// See trax_oracle.h for scope and the behaviours it fixes.  Build with
// -ffp-contract=off -fno-fast-math so every FP32 expression is evaluated
// exactly as written, in the reference's order.
#include "trax_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ofp32.h"

#define PI O_PI_F

static void *xcalloc(uint64_t n, uint64_t sz) {
  if (n == 0) n = 1;
  return calloc(n, sz);
}

// ---------------------------------------------------------------- prefix sum
// PrefixSum.h:55-245 / PrefixSum.cpp:60-126: a recursive Blelloch exclusive
// scan of uint32 values.  uint32 addition is associative mod 2^32, so the
// work-group tree gives exactly the sequential exclusive scan.
void oracle_scan_u32(uint32_t *data, uint64_t n) {
  uint32_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t v = data[i];
    data[i] = acc;
    acc += v;
  }
}

// ---------------------------------------------------------------- grid build
// GridBuilder.h:80-84 (and :239-241 in the store kernel).
uint32_t oracle_cell_of(const oracle_cfg *c, float x, float y, float z, float *phi_out) {
  uint32_t zs = o_f2u(floorf((z - c->min_z) / c->sector_size_z));
  float phi = o_atan2f(y, x);
  uint32_t ps = o_f2u(floorf((phi - c->min_phi) / c->sector_size_phi));
  if (zs > c->n_z) zs = c->n_z;  // defined clamp; in-range hits are unaffected
  if (ps > c->n_phi) ps = c->n_phi;
  if (phi_out) *phi_out = phi;
  return zs * (c->n_phi + 1) + ps;
}

static int grid_build(const oracle_cfg *c, const oracle_input *in, oracle_result *r) {
  const uint32_t E = c->n_events, L = c->n_layers;
  const uint64_t C = (uint64_t)(c->n_z + 1) * (c->n_phi + 1);
  const uint64_t G = (uint64_t)E * L * C;
  r->grid_size = G;
  r->grid = (uint32_t *)xcalloc(G + 1, 4);
  uint32_t *written = (uint32_t *)xcalloc(G, 4);
  const uint32_t n = in->n_hits;
  r->n_hits = n;
  r->x = (float *)xcalloc(n, 4);
  r->y = (float *)xcalloc(n, 4);
  r->z = (float *)xcalloc(n, 4);
  r->layer = (uint32_t *)xcalloc(n, 4);
  r->det_id = (uint32_t *)xcalloc(n, 4);
  r->hit_id = (uint32_t *)xcalloc(n, 4);
  r->evt = (uint32_t *)xcalloc(n, 4);
  r->orig_index = (uint32_t *)xcalloc(n, 4);
  if (!r->grid || !written || !r->x || !r->y || !r->z || !r->layer || !r->det_id ||
      !r->hit_id || !r->evt || !r->orig_index) {
    free(written);
    return -1;
  }
  // gridCount (GridBuilder.h:62-94): per (event, layer) histogram
  for (uint32_t e = 0; e < E; ++e)
    for (uint32_t l = 0; l < L; ++l) {
      uint32_t nh = in->layer_nhits[e * L + l];
      uint32_t off = in->event_offset[e] + in->layer_offset[e * L + l];
      uint32_t *g = r->grid + ((uint64_t)e * L + l) * C;
      for (uint32_t i = 0; i < nh; ++i) {
        uint32_t cell = oracle_cell_of(c, in->x[off + i], in->y[off + i], in->z[off + i], 0);
        g[cell] += 1;
      }
    }
  // PrefixSum over the whole grid (GridBuilder.cpp:58)
  oracle_scan_u32(r->grid, G);
  // gridStore (GridBuilder.h:191-253): w = grid[cell] + rank; rank in input order
  for (uint32_t e = 0; e < E; ++e)
    for (uint32_t l = 0; l < L; ++l) {
      uint32_t nh = in->layer_nhits[e * L + l];
      uint32_t off = in->event_offset[e] + in->layer_offset[e * L + l];
      uint64_t gb = ((uint64_t)e * L + l) * C;
      for (uint32_t i = off; i < off + nh; ++i) {
        uint32_t cell = oracle_cell_of(c, in->x[i], in->y[i], in->z[i], 0);
        uint32_t w = r->grid[gb + cell] + written[gb + cell]++;
        if (w >= n) continue;  // inconsistent layout; the reference would write out of bounds
        r->x[w] = in->x[i];
        r->y[w] = in->y[i];
        r->z[w] = in->z[i];
        r->layer[w] = in->layer[i];
        r->det_id[w] = in->det_id[i];
        r->hit_id[w] = in->hit_id[i];
        r->evt[w] = in->evt[i];
        r->orig_index[w] = i;
      }
    }
  free(written);
  return 0;
}

// --------------------------------------------------------------------- pairs
typedef struct {
  float zLow, zHigh, phiLow, phiHigh;
  int wrap;
  uint32_t zLowSector, zHighSector, phiLowSector, phiHighSector;
} window_t;

// pairCount window of one outer hit, PairGeneratorBeamspot.h:110-162.
static void pair_window(const oracle_cfg *c, uint32_t p, float X, float Y, float Z,
                        float minR1, float maxR1, window_t *w) {
  const float z0 = c->z0[p], tip = c->tip_cut[p], sigmaZ = c->sigma_z[p];
  // The kernel's gSigmaPhi argument receives the uint pairSpreadPhi column
  // (PairGeneratorBeamspot.cpp:55), so its bits are read as a float.
  float sigmaPhi;
  memcpy(&sigmaPhi, &c->pair_spread_phi[p], 4);
  const float R = c->min_radius_curvature;

  float signRadius = o_sign(Y);
  float r = signRadius * sqrtf(X * X + Y * Y);
  float cotThetaHigh = (Z - z0) / r;
  float cotThetaLow = (Z + z0) / r;
  float tmp = signRadius * maxR1 * cotThetaHigh + z0;
  float zHigh = signRadius * minR1 * cotThetaHigh + z0;
  zHigh = o_b2f(tmp < zHigh) * zHigh + o_b2f(tmp > zHigh) * tmp;
  tmp = signRadius * maxR1 * cotThetaLow - z0;
  float zLow = signRadius * minR1 * cotThetaLow - z0;
  zLow = o_b2f(tmp < zLow) * tmp + o_b2f(tmp > zLow) * zLow;
  zLow -= sigmaZ;
  zHigh += sigmaZ;
  int32_t zl = o_f2i(floorf((zLow - c->min_z) / c->sector_size_z));
  uint32_t zLowSector = (uint32_t)(zl > 0 ? zl : 0);
  uint32_t zh = o_f2u(floorf((zHigh - c->min_z) / c->sector_size_z)) + 1u;
  uint32_t zHighSector = zh < c->n_z ? zh : c->n_z;

  float phi = o_atan2f(Y, X);
  tmp = fabsf(o_acosf(r / (2 * R)) - o_acosf(maxR1 / (2 * R)));
  float dPhi = fabsf(o_acosf(r / (2 * R)) - o_acosf(minR1 / (2 * R)));
  dPhi = o_b2f(tmp < dPhi) * dPhi + o_b2f(tmp > dPhi) * tmp;
  tmp = o_atanf(tip * (r - minR1)) / (r * minR1);
  float phiHigh = o_atanf(tip * (r - maxR1)) / (r * maxR1);
  dPhi += tmp > phiHigh ? tmp : phiHigh;
  float phiLow = phi - dPhi - sigmaPhi;
  phiHigh = phi + dPhi + sigmaPhi;
  int wrap = phiLow < -PI || phiHigh > PI || phiLow > PI || phiHigh < -PI;
  phiLow -= (float)((phiLow > PI) * 2) * PI;
  phiLow += (float)((phiLow < -PI) * 2) * PI;
  phiHigh -= (float)((phiHigh > PI) * 2) * PI;
  phiHigh += (float)((phiHigh < -PI) * 2) * PI;
  uint32_t phiLowSector = o_f2u(floorf((phiLow - c->min_phi) / c->sector_size_phi));
  uint32_t phiHighSector = o_f2u(floorf((phiHigh - c->min_phi) / c->sector_size_phi)) + 1u;
  // defined clamps: only reached for windows wider than 2*pi
  if (phiLowSector > c->n_phi) phiLowSector = c->n_phi;
  if (phiHighSector > c->n_phi + 1) phiHighSector = c->n_phi + 1;

  w->zLow = zLow;
  w->zHigh = zHigh;
  w->phiLow = phiLow;
  w->phiHigh = phiHigh;
  w->wrap = wrap;
  w->zLowSector = zLowSector;
  w->zHighSector = zHighSector;
  w->phiLowSector = phiLowSector;
  w->phiHighSector = phiHighSector;
}

static inline int phi_ok(const window_t *w, float actPhi) {
  return (!w->wrap && (w->phiLow <= actPhi && actPhi <= w->phiHigh)) ||
         (w->wrap && ((w->phiLow <= actPhi && actPhi <= PI) ||
                      (-PI <= actPhi && actPhi <= w->phiHigh)));
}

static int pairs_stage(const oracle_cfg *c, oracle_result *r) {
  const uint32_t E = c->n_events, L = c->n_layers, P = c->n_triplets, T = c->threads;
  const uint32_t nPhi = c->n_phi;
  const uint64_t C = (uint64_t)(c->n_z + 1) * (nPhi + 1);
  const uint64_t nCounts = (uint64_t)E * P * T + 1;
  uint32_t *prefix = (uint32_t *)xcalloc(nCounts, 4);
  r->pair_offsets = (uint32_t *)xcalloc((uint64_t)E * P + 1, 4);
  if (!prefix || !r->pair_offsets) return -1;
  // one bitmap per (event, layer pair), n1*n2 bits rounded to 32
  // (PairGeneratorBeamspot.cpp:13-34); kept per block instead of one array.
  uint32_t **bitmap = (uint32_t **)xcalloc((uint64_t)E * P, sizeof(uint32_t *));
  uint64_t raw_total = 0;
  for (uint32_t e = 0; e < E; ++e)
    for (uint32_t p = 0; p < P; ++p) {
      const uint64_t k = (uint64_t)e * P + p;
      const uint32_t *g2 = r->grid + ((uint64_t)e * L + (c->layer2[p] - 1)) * C;
      const uint32_t *g1 = r->grid + ((uint64_t)e * L + (c->layer1[p] - 1)) * C;
      const uint32_t layer2Offset = g2[0], end = g2[C - 1], nHits2 = end - layer2Offset;
      const uint32_t offset = g1[0], nHits1 = g1[C - 1] - g1[0];
      const uint64_t nbits = (uint64_t)nHits1 * nHits2;
      bitmap[k] = (uint32_t *)xcalloc((nbits + 31) / 32 + 1, 4);
      if (!bitmap[k]) return -1;
      const float minR1 = c->min_radius[c->layer1[p] - 1];
      const float maxR1 = c->max_radius[c->layer1[p] - 1];
      // pairCount (PairGeneratorBeamspot.h:108-234): work-item t handles
      // outer hits layer2Offset + t, + T, ...
      for (uint32_t i = layer2Offset; i < end; ++i) {
        const uint32_t t = (i - layer2Offset) % T;
        window_t w;
        pair_window(c, p, r->x[i], r->y[i], r->z[i], minR1, maxR1, &w);
        uint32_t nFound = 0;
        for (uint32_t zs = w.zLowSector; zs < w.zHighSector; ++zs) {
          const uint32_t rowBase = zs * (nPhi + 1);
          const uint32_t zSectorStart = g1[rowBase];
          const uint32_t zSectorEnd = g1[rowBase + nPhi];  // (zs+1)(nPhi+1)-1
          const uint32_t zSectorLength = zSectorEnd - zSectorStart;
          uint32_t j = g1[rowBase + w.phiLowSector];
          const uint32_t end2 = (uint32_t)w.wrap * zSectorEnd + g1[rowBase + w.phiHighSector] -
                                (uint32_t)w.wrap * zSectorStart;
          for (; j < end2; ++j) {
            const uint32_t index = j - (uint32_t)(j >= zSectorEnd) * zSectorLength;
            int valid = w.zLow <= r->z[index] && r->z[index] <= w.zHigh;
            float actPhi = o_atan2f(r->y[index], r->x[index]);
            valid = valid && phi_ok(&w, actPhi);
            nFound += (uint32_t)valid;
            if (valid) {
              uint32_t bit = (index - offset) * nHits2 + i - layer2Offset;
              bitmap[k][bit / 32] |= 1u << (bit % 32);
            }
          }
        }
        prefix[k * T + t] += nFound;
        raw_total += nFound;
      }
    }
  oracle_scan_u32(prefix, nCounts);
  const uint32_t nPairs = prefix[nCounts - 1];
  r->n_pairs = nPairs;
  r->pairs = (uint32_t *)xcalloc(2 * (uint64_t)nPairs, 4);
  if (!r->pairs) return -1;
  memset(r->pairs, 0xFF, 8 * (uint64_t)nPairs);
  uint64_t written = 0;
  // pairStore (PairGeneratorBeamspot.h:459-521)
  for (uint32_t e = 0; e < E; ++e)
    for (uint32_t p = 0; p < P; ++p) {
      const uint64_t k = (uint64_t)e * P + p;
      const uint32_t *g2 = r->grid + ((uint64_t)e * L + (c->layer2[p] - 1)) * C;
      const uint32_t *g1 = r->grid + ((uint64_t)e * L + (c->layer1[p] - 1)) * C;
      const uint32_t layer2Offset = g2[0], end = g2[C - 1], nHits2 = end - layer2Offset;
      const uint32_t offset = g1[0], nHits1 = g1[C - 1] - g1[0];
      for (uint32_t t = 0; t < T; ++t) {
        uint32_t pos = prefix[k * T + t];
        const uint32_t next = prefix[k * T + t + 1];
        for (uint32_t i = layer2Offset + t; i < end; i += T)
          for (uint32_t j = 0; j < nHits1 && pos < next; ++j) {
            uint32_t bit = j * nHits2 + i - layer2Offset;
            int valid = (bitmap[k][bit / 32] >> (bit % 32)) & 1u;
            if (valid) {
              r->pairs[2 * (uint64_t)pos] = j + offset;
              r->pairs[2 * (uint64_t)pos + 1] = i;
              ++written;
            }
            pos += (uint32_t)valid;
          }
      }
      r->pair_offsets[k + 1] = prefix[k * T + T];
    }
  r->n_pairs_unwritten = (uint32_t)(nPairs - written);
  r->n_pairs_raw_overcount = (uint32_t)(raw_total - written);
  for (uint64_t k = 0; k < (uint64_t)E * P; ++k) free(bitmap[k]);
  free(bitmap);
  free(prefix);
  return 0;
}

// ------------------------------------------------------------ sector pairs
// PairGeneratorSector.h:43-234, .cpp:5-151: every inner-layer hit with every
// outer-layer hit in the grid cells of its (z row +- pairSpreadZ) x (phi
// cell +- pairSpreadPhi) window, no geometric cut; n1*n2 bitmap per (event,
// layer pair), per-work-item visit counts, scan, bitmap-ordered store.

typedef struct {
  uint32_t zLowSector, zHighSector, phiLowSector, phiHighSector;
  int wrap;
} sector_window_t;

// pairCount window (PairGeneratorSector.h:100-115), with the OpenCL integer
// types: zLowSector - spreadZ is unsigned and wraps below row 0 (the row loop
// then runs zero times); the phi sector is an int from float arithmetic.
static void sector_window(const oracle_cfg *c, uint32_t p, float X, float Y, float Z,
                          sector_window_t *w) {
  const uint32_t spreadZ = c->pair_spread_z[p], spreadPhi = c->pair_spread_phi[p];
  const uint32_t nZ = c->n_z, nPhi = c->n_phi;
  uint32_t zLowSector = o_f2u(floorf((Z - c->min_z) / c->sector_size_z));
  uint32_t zHighSector = zLowSector + spreadZ + 1u;
  zHighSector = zHighSector < nZ ? zHighSector : nZ;  // min(uint, uint)
  zLowSector = zLowSector - spreadZ;                  // max(0u, .) of a uint
  const float phi = o_atan2f(Y, X);
  int32_t phiLowSector = o_f2i(floorf((phi - c->min_phi) / c->sector_size_phi) - (float)spreadPhi);
  uint32_t phiHighSector = (uint32_t)phiLowSector + 2u * spreadPhi + 1u;
  const int wrap = phiLowSector < 0 || phiHighSector > nPhi + 1u;
  phiLowSector = (int32_t)((uint32_t)phiLowSector + (uint32_t)(phiLowSector < 0) * nPhi);
  phiHighSector -= (uint32_t)(phiHighSector > nPhi + 1u) * nPhi;
  w->zLowSector = zLowSector;
  w->zHighSector = zHighSector;
  w->phiLowSector = (uint32_t)phiLowSector;
  w->phiHighSector = phiHighSector;
  w->wrap = wrap;
}

static int sector_pairs_stage(const oracle_cfg *c, oracle_result *r) {
  const uint32_t E = c->n_events, L = c->n_layers, P = c->n_triplets, T = c->threads;
  const uint32_t nPhi = c->n_phi;
  const uint64_t C = (uint64_t)(c->n_z + 1) * (nPhi + 1);
  const uint64_t nCounts = (uint64_t)E * P * T + 1;
  uint32_t *prefix = (uint32_t *)xcalloc(nCounts, 4);
  r->pair_offsets = (uint32_t *)xcalloc((uint64_t)E * P + 1, 4);
  if (!prefix || !r->pair_offsets) return -1;
  uint32_t **bitmap = (uint32_t **)xcalloc((uint64_t)E * P, sizeof(uint32_t *));
  if (!bitmap) return -1;
  uint64_t raw_total = 0;
  for (uint32_t e = 0; e < E; ++e)
    for (uint32_t p = 0; p < P; ++p) {
      const uint64_t k = (uint64_t)e * P + p;
      const uint32_t *g1 = r->grid + ((uint64_t)e * L + (c->layer1[p] - 1)) * C;
      const uint32_t *g2 = r->grid + ((uint64_t)e * L + (c->layer2[p] - 1)) * C;
      // inner hits up to the last cell's start; outer likewise (:67-81)
      const uint32_t layer1Offset = g1[0], end = g1[C - 1];
      const uint32_t offset = g2[0], nHits2 = g2[C - 1] - g2[0];
      const uint64_t nbits = (uint64_t)(end - layer1Offset) * nHits2;
      bitmap[k] = (uint32_t *)xcalloc((nbits + 31) / 32 + 1, 4);
      if (!bitmap[k]) return -1;
      // work-item t handles inner hits layer1Offset + t, + T, ... (:95-146)
      for (uint32_t i = layer1Offset; i < end; ++i) {
        const uint32_t t = (i - layer1Offset) % T;
        sector_window_t w;
        sector_window(c, p, r->x[i], r->y[i], r->z[i], &w);
        uint32_t nFound = 0;
        for (uint32_t zSector = w.zLowSector; zSector < w.zHighSector; ++zSector) {
          const uint32_t row = zSector * (nPhi + 1);
          const uint32_t zSectorStart = g2[row];
          const uint32_t zSectorEnd = g2[row + nPhi];  // (zSector+1)(nPhi+1)-1
          const uint32_t zSectorLength = zSectorEnd - zSectorStart;
          uint32_t j = g2[row + w.phiLowSector];
          const uint32_t end2 = (uint32_t)w.wrap * zSectorEnd + g2[row + w.phiHighSector] -
                                (uint32_t)w.wrap * zSectorStart;
          for (; j < end2; ++j) {
            ++nFound;
            const uint32_t index =
                (i - layer1Offset) * nHits2 + j - (uint32_t)(j >= zSectorEnd) * zSectorLength - offset;
            bitmap[k][index / 32] |= 1u << (index % 32);
          }
        }
        prefix[k * T + t] += nFound;
        raw_total += nFound;
      }
    }
  oracle_scan_u32(prefix, nCounts);
  const uint32_t nPairs = prefix[nCounts - 1];
  r->n_pairs = nPairs;
  r->pairs = (uint32_t *)xcalloc(2 * (uint64_t)nPairs + 2, 4);
  if (!r->pairs) return -1;
  memset(r->pairs, 0xFF, 8 * (uint64_t)nPairs);
  uint64_t written = 0;
  // pairStore (PairGeneratorSector.h:157-234): per work item, every inner hit
  // of its stride tests all n2 bits and writes (inner, outer) while pos < next
  for (uint32_t e = 0; e < E; ++e)
    for (uint32_t p = 0; p < P; ++p) {
      const uint64_t k = (uint64_t)e * P + p;
      const uint32_t *g1 = r->grid + ((uint64_t)e * L + (c->layer1[p] - 1)) * C;
      const uint32_t *g2 = r->grid + ((uint64_t)e * L + (c->layer2[p] - 1)) * C;
      const uint32_t layer1Offset = g1[0], end = g1[C - 1];
      const uint32_t offset = g2[0], nHits2 = g2[C - 1] - g2[0];
      for (uint32_t t = 0; t < T; ++t) {
        uint32_t pos = prefix[k * T + t];
        const uint32_t next = prefix[k * T + t + 1];
        for (uint32_t i = layer1Offset + t; i < end; i += T)
          for (uint32_t j = 0; j < nHits2 && pos < next; ++j) {
            const uint32_t index = (i - layer1Offset) * nHits2 + j;
            const int valid = (bitmap[k][index / 32] >> (index % 32)) & 1u;
            if (valid) {
              r->pairs[2 * (uint64_t)pos] = i;
              r->pairs[2 * (uint64_t)pos + 1] = offset + j;
              ++written;
            }
            pos += (uint32_t)valid;
          }
      }
      r->pair_offsets[k + 1] = prefix[k * T + T];
    }
  r->n_pairs_unwritten = (uint32_t)(nPairs - written);
  r->n_pairs_raw_overcount = (uint32_t)(raw_total - written);
  for (uint64_t k = 0; k < (uint64_t)E * P; ++k) free(bitmap[k]);
  free(bitmap);
  free(prefix);
  return 0;
}

// ---------------------------------------------------------------- prediction
typedef struct {
  float zLow, zHigh, phiLow, phiHigh;
  int wrap;
  uint32_t zLowSector, zHighSector, phiLowSector, phiHighSector;
} pwindow_t;

// predictCount (TripletThetaPhiPredictor.h:101-166) when store == 0,
// predictStore (:310-375) when store == 1.  They differ in rOrigin
// (:117-118 vs :326-327) and in the zHigh margin (:130 sigmaPhi vs :339 sigmaZ).
static void predict_window(const oracle_cfg *c, uint32_t lt, int store, float p1x, float p1y,
                           float p1z, float p2x, float p2y, float p2z, float minR, float maxR,
                           pwindow_t *w) {
  const float sigmaZ = c->sigma_z[lt], sigmaPhi = c->sigma_phi[lt];
  const float R = c->min_radius_curvature;
  float tip2 = p1y * (p2x - p1x) - p1x * (p2y - p1y);
  tip2 *= tip2;
  tip2 /= ((p2x - p1x) * (p2x - p1x) + (p2y - p1y) * (p2y - p1y));
  float rOrigin;
  if (!store) {
    rOrigin = p1x * p1x + p1y * p1y;
    rOrigin = sqrtf(rOrigin - tip2);
  } else {
    rOrigin = sqrtf(p1x * p1x + p1y * p1y);
    rOrigin = sqrtf(rOrigin * rOrigin - tip2);
  }
  float cotTheta = (p2z - p1z) / (sqrtf(p2x * p2x + p2y * p2y - tip2) - rOrigin);
  float tmp = p1z + (sqrtf(maxR * maxR - tip2) - rOrigin) * cotTheta;
  float zLow = p1z + (sqrtf(minR * minR - tip2) - rOrigin) * cotTheta;
  float zHigh = o_b2f(tmp > zLow) * tmp + o_b2f(tmp < zLow) * zLow;
  zLow = o_b2f(tmp > zLow) * zLow + o_b2f(tmp < zLow) * tmp;
  zLow -= sigmaZ;
  zHigh += store ? sigmaZ : sigmaPhi;
  int32_t zl = o_f2i(floorf((zLow - c->min_z) / c->sector_size_z));
  uint32_t zLowSector = (uint32_t)(zl > 0 ? zl : 0);
  uint32_t zh = o_f2u(floorf((zHigh - c->min_z) / c->sector_size_z)) + 1u;
  uint32_t zHighSector = zh < c->n_z ? zh : c->n_z;

  float phi = o_atan2f(p2y, p2x);
  tmp = o_atan2f(p1y, p1x);
  float dPhi = phi - tmp;
  dPhi += (dPhi > PI) ? -2 * PI : (dPhi < -PI) ? 2 * PI : 0;
  dPhi = fabsf(dPhi);
  float dHits = sqrtf((p2x - p1x) * (p2x - p1x) + (p2y - p1y) * (p2y - p1y));
  tmp = fabsf(o_acosf(dHits / (2 * R)) - o_acosf(maxR / (2 * R)));
  float phiLow = fabsf(o_acosf(dHits / (2 * R)) - o_acosf(minR / (2 * R)));
  dPhi = o_maxf(dPhi, o_maxf(tmp, phiLow));
  float phiHigh = phi + dPhi + sigmaPhi;
  phiLow = phi - dPhi - sigmaPhi;
  int wrap = phiLow < -PI || phiHigh > PI || phiLow > PI || phiHigh < -PI;
  phiLow -= (float)((phiLow > PI) * 2) * PI;
  phiLow += (float)((phiLow < -PI) * 2) * PI;
  phiHigh -= (float)((phiHigh > PI) * 2) * PI;
  phiHigh += (float)((phiHigh < -PI) * 2) * PI;
  uint32_t phiLowSector = o_f2u(floorf((phiLow - c->min_phi) / c->sector_size_phi));
  uint32_t ph = o_f2u(floorf((phiHigh - c->min_phi) / c->sector_size_phi)) + 1u;
  uint32_t phiHighSector = ph < c->n_phi ? ph : c->n_phi;
  if (phiLowSector > c->n_phi) phiLowSector = c->n_phi;  // defined clamp (window > 2*pi)

  w->zLow = zLow;
  w->zHigh = zHigh;
  w->phiLow = phiLow;
  w->phiHigh = phiHigh;
  w->wrap = wrap;
  w->zLowSector = zLowSector;
  w->zHighSector = zHighSector;
  w->phiLowSector = phiLowSector;
  w->phiHighSector = phiHighSector;
}

static int predict_stage(const oracle_cfg *c, oracle_result *r) {
  const uint32_t E = c->n_events, L = c->n_layers, P = c->n_triplets, nPhi = c->n_phi;
  const uint64_t C = (uint64_t)(c->n_z + 1) * (nPhi + 1);
  const uint32_t nPairs = r->n_pairs;
  uint32_t *prefix = (uint32_t *)xcalloc((uint64_t)nPairs + 1, 4);
  r->cand_offsets = (uint32_t *)xcalloc((uint64_t)E * P + 1, 4);
  if (!prefix || !r->cand_offsets) return -1;
  for (int store = 0; store < 2; ++store) {
    if (store) {
      oracle_scan_u32(prefix, (uint64_t)nPairs + 1);
      r->n_cands = prefix[nPairs];
      r->cands = (uint32_t *)xcalloc(2 * (uint64_t)r->n_cands, 4);
      if (!r->cands) return -1;
      memset(r->cands, 0xFF, 8 * (uint64_t)r->n_cands);
    }
    uint64_t written = 0;
    for (uint32_t gid = 0; gid < nPairs; ++gid) {
      const uint32_t firstHit = r->pairs[2 * (uint64_t)gid];
      const uint32_t secondHit = r->pairs[2 * (uint64_t)gid + 1];
      if (firstHit == ORACLE_UNWRITTEN || secondHit == ORACLE_UNWRITTEN) continue;
      const uint32_t event = r->evt[firstHit];
      const uint32_t lt = r->layer[firstHit] - 1;  // :79 (innermost layer - 1)
      if (event >= E || lt >= P) continue;
      const uint32_t layer = c->layer3[lt] - 1;
      const uint32_t *g3 = r->grid + ((uint64_t)event * L + layer) * C;
      pwindow_t w;
      predict_window(c, lt, store, r->x[firstHit], r->y[firstHit], r->z[firstHit],
                     r->x[secondHit], r->y[secondHit], r->z[secondHit], c->min_radius[layer],
                     c->max_radius[layer], &w);
      uint32_t nFound = 0;
      uint32_t pos = store ? prefix[gid] : 0;
      const uint32_t next = store ? prefix[gid + 1] : 0;
      for (uint32_t zs = w.zLowSector; zs < w.zHighSector; ++zs) {
        const uint32_t rowBase = zs * (nPhi + 1);
        const uint32_t zSectorStart = g3[rowBase];
        const uint32_t zSectorEnd = g3[rowBase + nPhi + 1];  // (zs+1)(nPhi+1)
        const uint32_t zSectorLength = zSectorEnd - zSectorStart;
        uint32_t j = g3[rowBase + w.phiLowSector];
        const uint32_t end2 = (uint32_t)w.wrap * zSectorEnd + g3[rowBase + w.phiHighSector] -
                              (uint32_t)w.wrap * zSectorStart;
        for (; j < end2 && (!store || pos < next); ++j) {
          const uint32_t index = j - (uint32_t)(j >= zSectorEnd) * zSectorLength;
          int valid = w.zLow <= r->z[index] && r->z[index] <= w.zHigh;
          float actPhi = o_atan2f(r->y[index], r->x[index]);
          valid = valid && ((!w.wrap && (w.phiLow <= actPhi && actPhi <= w.phiHigh)) ||
                            (w.wrap && ((w.phiLow <= actPhi && actPhi <= PI) ||
                                        (-PI <= actPhi && actPhi <= w.phiHigh))));
          if (!store) {
            nFound += (uint32_t)valid;
          } else {
            if (valid) {
              r->cands[2 * (uint64_t)pos] = gid;
              r->cands[2 * (uint64_t)pos + 1] = index;
              ++written;
            }
            pos += (uint32_t)valid;
          }
        }
      }
      if (!store) {
        prefix[gid] = nFound;
      } else {
        // per-(event, layer triplet) end offsets (:459-469)
        int last = 1;
        if (gid < nPairs - 1) {
          const uint32_t nextHit = r->pairs[2 * (uint64_t)(gid + 1)];
          if (nextHit != ORACLE_UNWRITTEN)
            last = (r->layer[nextHit] - 1 != lt) || (r->evt[nextHit] != event);
        }
        if (last) r->cand_offsets[(uint64_t)event * P + lt + 1] = next;
      }
    }
    if (store) r->n_cands_unwritten = (uint32_t)(r->n_cands - written);
  }
  // predictOffsetMonotonizeStore (:483-497), as the sequential fill it intends
  for (uint64_t g = 1; g < (uint64_t)E * P + 1; ++g)
    if (r->cand_offsets[g] == 0) r->cand_offsets[g] = r->cand_offsets[g - 1];
  free(prefix);
  return 0;
}

// -------------------------------------------------------------------- filter
typedef struct {
  float x, y, z;
} f3;

static inline f3 f3_normalize(f3 v) {
  float len = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  if (len == 0.0f) return v;
  f3 o = {v.x / len, v.y / len, v.z / len};
  return o;
}

// filterCount predicate, TripletThetaPhiFilter.h:68-147.
static int filter_pass_ex(const oracle_cfg *c, uint32_t lt, float x1, float y1, float z1, float x2,
                          float y2, float z2, float x3, float y3, float z3, float *info) {
  const float dThetaCut = c->dtheta_cut[lt], dPhiCut = c->dphi_cut[lt], tipCut = c->tip_cut[lt];
  const float minRadius = c->min_radius_curvature;
  int valid = 1;
  float angle1 = o_atan2f(sqrtf((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1)), (z2 - z1));
  float angle2 = o_atan2f(sqrtf((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2)), (z3 - z2));
  float delta = fabsf(angle2 / angle1);
  valid = valid && (1 - dThetaCut <= delta && delta <= 1 + dThetaCut);
  if (info) info[0] = delta;

  angle1 = o_atan2f((y2 - y1), (x2 - x1));
  angle2 = o_atan2f((y3 - y2), (x3 - x2));
  delta = angle2 - angle1;
  delta += (delta > PI) ? -2 * PI : (delta < -PI) ? 2 * PI : 0;
  valid = valid && (fabsf(delta) <= dPhiCut);
  if (info) info[1] = fabsf(delta);

  f3 pP1 = {x1, y1, x1 * x1 + y1 * y1};
  f3 pP2 = {x2, y2, x2 * x2 + y2 * y2};
  f3 pP3 = {x3, y3, x3 * x3 + y3 * y3};
  f3 a = {pP2.x - pP1.x, pP2.y - pP1.y, pP2.z - pP1.z};
  f3 b = {pP3.x - pP1.x, pP3.y - pP1.y, pP3.z - pP1.z};
  f3 n = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  n = f3_normalize(n);
  float cOx = (-n.x) / (2 * n.z);
  float cOy = (-n.y) / (2 * n.z);
  float cc = -(n.x * pP1.x + n.y * pP1.y + n.z * pP1.z);
  float cR = sqrtf((1 - n.z * n.z - 4 * cc * n.z) / (4 * n.z * n.z));
  float vx = -cOx, vy = -cOy;
  float vlen = sqrtf(vx * vx + vy * vy);
  if (vlen != 0.0f) {
    vx = vx / vlen;
    vy = vy / vlen;
  }
  float pCAx = cOx + cR * vx;
  float pCAy = cOy + cR * vy;
  float tip = pCAx * pCAx + pCAy * pCAy;
  valid = valid && (tip <= (tipCut * tipCut));
  valid = valid && (cR >= minRadius);
  if (info) {
    info[2] = tip;
    info[3] = cR;
    info[4] = (float)valid;
  }
  return valid;
}

static int filter_pass(const oracle_cfg *c, uint32_t lt, float x1, float y1, float z1, float x2,
                       float y2, float z2, float x3, float y3, float z3) {
  return filter_pass_ex(c, lt, x1, y1, z1, x2, y2, z2, x3, y3, z3, 0);
}

static int filter_stage(const oracle_cfg *c, oracle_result *r) {
  const uint32_t E = c->n_events, P = c->n_triplets;
  const uint32_t nC = r->n_cands;
  uint8_t *keep = (uint8_t *)xcalloc(nC, 1);
  r->tracklet_offsets = (uint32_t *)xcalloc((uint64_t)E * P + 1, 4);
  if (!keep || !r->tracklet_offsets) return -1;
  uint32_t nT = 0;
  for (uint32_t gid = 0; gid < nC; ++gid) {
    const uint32_t pi = r->cands[2 * (uint64_t)gid], h3 = r->cands[2 * (uint64_t)gid + 1];
    if (pi == ORACLE_UNWRITTEN || h3 == ORACLE_UNWRITTEN) continue;
    const uint32_t h1 = r->pairs[2 * (uint64_t)pi], h2 = r->pairs[2 * (uint64_t)pi + 1];
    const uint32_t lt = r->layer[h1] - 1;  // :73
    if (lt >= P) continue;
    keep[gid] = (uint8_t)filter_pass(c, lt, r->x[h1], r->y[h1], r->z[h1], r->x[h2], r->y[h2],
                                     r->z[h2], r->x[h3], r->y[h3], r->z[h3]);
    nT += keep[gid];
  }
  // popcount -> prefix sum -> store in candidate order (:170-252)
  r->n_tracklets = nT;
  r->t1 = (uint32_t *)xcalloc(nT, 4);
  r->t2 = (uint32_t *)xcalloc(nT, 4);
  r->t3 = (uint32_t *)xcalloc(nT, 4);
  if (!r->t1 || !r->t2 || !r->t3) return -1;
  uint32_t pos = 0;
  for (uint32_t gid = 0; gid < nC; ++gid)
    if (keep[gid]) {
      const uint32_t pi = r->cands[2 * (uint64_t)gid];
      r->t1[pos] = r->pairs[2 * (uint64_t)pi];
      r->t2[pos] = r->pairs[2 * (uint64_t)pi + 1];
      r->t3[pos] = r->cands[2 * (uint64_t)gid + 1];
      ++pos;
    }
  // filterOffsetStore (:270-292) + the intended monotone fill (:309-312)
  for (uint32_t gid = 0; gid < nT; ++gid) {
    const uint32_t lt = r->layer[r->t1[gid]] - 1, ev = r->evt[r->t1[gid]];
    int last = 1;
    if (gid < nT - 1) {
      const uint32_t h = r->t1[gid + 1];
      last = (r->layer[h] - 1 != lt) || (r->evt[h] != ev);
    }
    if (last && ev < E && lt < P) r->tracklet_offsets[(uint64_t)ev * P + lt + 1] = gid + 1;
  }
  for (uint64_t g = 1; g < (uint64_t)E * P + 1; ++g)
    if (r->tracklet_offsets[g] == 0) r->tracklet_offsets[g] = r->tracklet_offsets[g - 1];
  free(keep);
  return 0;
}

// ------------------------------------------------------------------- driver
int oracle_run_chain(const oracle_cfg *c, const oracle_input *in, int stop_after,
                     oracle_result *out) {
  memset(out, 0, sizeof(*out));
  if (c->n_layers == 0 || c->n_layers > ORACLE_MAX_LAYERS || c->n_triplets > ORACLE_MAX_TRIPLETS ||
      c->threads == 0)
    return -2;
  // the sector window indexes inside its grid row only for 2 * spread <= nPhi
  if (c->pair_generator == ORACLE_PAIRS_SECTOR)
    for (uint32_t p = 0; p < c->n_triplets; ++p)
      if (2u * c->pair_spread_phi[p] > c->n_phi) return -2;
  for (uint32_t p = 0; p < c->n_triplets; ++p)
    if (c->layer1[p] < 1 || c->layer1[p] > c->n_layers || c->layer2[p] < 1 ||
        c->layer2[p] > c->n_layers || c->layer3[p] < 1 || c->layer3[p] > c->n_layers)
      return -2;
  if (grid_build(c, in, out)) return -1;
  if (stop_after < 1) return 0;
  if ((c->pair_generator == ORACLE_PAIRS_SECTOR ? sector_pairs_stage(c, out) : pairs_stage(c, out)))
    return -1;
  if (stop_after < 2) return 0;
  if (predict_stage(c, out)) return -1;
  if (stop_after < 3) return 0;
  if (filter_stage(c, out)) return -1;
  return 0;
}

void oracle_free_result(oracle_result *r) {
  free(r->x);
  free(r->y);
  free(r->z);
  free(r->layer);
  free(r->det_id);
  free(r->hit_id);
  free(r->evt);
  free(r->orig_index);
  free(r->grid);
  free(r->pairs);
  free(r->pair_offsets);
  free(r->cands);
  free(r->cand_offsets);
  free(r->t1);
  free(r->t2);
  free(r->t3);
  free(r->tracklet_offsets);
  memset(r, 0, sizeof(*r));
}

void oracle_eval_f32(int op, const float *a, const float *b, float *out, uint64_t n) {
  for (uint64_t i = 0; i < n; ++i) {
    switch (op) {
      case 0: out[i] = o_atan2f(a[i], b[i]); break;
      case 1: out[i] = o_atanf(a[i]); break;
      case 2: out[i] = o_acosf(a[i]); break;
      case 3: out[i] = o_sign(a[i]); break;
      case 4: out[i] = a[i] / b[i]; break;
      case 5: out[i] = sqrtf(a[i]); break;
      default: out[i] = 0.0f;
    }
  }
}

void oracle_sector_window(const oracle_cfg *c, uint32_t p, float x, float y, float z, uint32_t *out) {
  sector_window_t w;
  sector_window(c, p, x, y, z, &w);
  out[0] = w.zLowSector;
  out[1] = w.zHighSector;
  out[2] = w.phiLowSector;
  out[3] = w.phiHighSector;
  out[4] = (uint32_t)w.wrap;
}

void oracle_pair_window(const oracle_cfg *c, uint32_t p, float x, float y, float z, float *out) {
  window_t w;
  pair_window(c, p, x, y, z, c->min_radius[c->layer1[p] - 1], c->max_radius[c->layer1[p] - 1], &w);
  out[0] = w.zLow;
  out[1] = w.zHigh;
  out[2] = w.phiLow;
  out[3] = w.phiHigh;
  out[4] = (float)w.wrap;
  out[5] = (float)w.zLowSector;
  out[6] = (float)w.zHighSector;
  out[7] = (float)w.phiLowSector;
  out[8] = (float)w.phiHighSector;
}

void oracle_predict_window(const oracle_cfg *c, uint32_t lt, int store, const float *p1,
                           const float *p2, float minR, float maxR, float *out) {
  pwindow_t w;
  predict_window(c, lt, store, p1[0], p1[1], p1[2], p2[0], p2[1], p2[2], minR, maxR, &w);
  out[0] = w.zLow;
  out[1] = w.zHigh;
  out[2] = w.phiLow;
  out[3] = w.phiHigh;
  out[4] = (float)w.wrap;
  out[5] = (float)w.zLowSector;
  out[6] = (float)w.zHighSector;
  out[7] = (float)w.phiLowSector;
  out[8] = (float)w.phiHighSector;
}

void oracle_filter_eval(const oracle_cfg *c, uint32_t lt, const float *h1, const float *h2,
                        const float *h3, float *out) {
  filter_pass_ex(c, lt, h1[0], h1[1], h1[2], h2[0], h2[1], h2[2], h3[0], h3[1], h3[2], out);
}

// TripletConnectivity (TripletConnectivity.cpp:21-71 selects the columns;
// the kernels are TripletConnectivity.h:56-108).  The loop is the reference's
// one work-item per following triplet over the base range [0, nb - 1], with
// the join the reference's known answers fix (TripletConnectivityTest.h:
// test_run, test_run_backward, test_run_multilayer, test_run_tight_multilayer):
// the wide kernel as printed indexes the base column with the following
// triplet's id and the following column with the base loop counter; that
// order gives 2 instead of the asserted 3 on test_run_multilayer and reads out
// of bounds when the two sets differ in size.
int oracle_connectivity(const uint32_t *bh1, const uint32_t *bh2, const uint32_t *bh3,
                        const uint32_t *bcon, uint64_t nb, const uint32_t *fh1, const uint32_t *fh2,
                        const uint32_t *fh3, uint32_t *fcon, uint64_t nf, int backward, int tight) {
  if (tight && backward) return -1;  // assert(!iterateBackwards), TripletConnectivity.cpp:56
  if (nb == 0 || nf == 0) return 0;
  uint32_t *w = (uint32_t *)malloc(nb * sizeof(uint32_t));
  if (!w) return -2;
  for (uint64_t i = 0; i < nb; ++i) w[i] = bcon[i] + 1u;
  for (uint64_t g = 0; g < nf; ++g) {
    uint32_t acc = fcon[g];
    for (uint64_t i = 0; i < nb; ++i) {
      int joined;
      if (tight)
        joined = bh2[i] == fh1[g] && bh3[i] == fh2[g];  // tripletConnectivityTight, :95-98
      else if (backward)
        joined = bh1[i] == fh3[g];
      else
        joined = bh3[i] == fh1[g];
      acc += (uint32_t)joined * w[i];
    }
    fcon[g] = acc;
  }
  free(w);
  return 0;
}
