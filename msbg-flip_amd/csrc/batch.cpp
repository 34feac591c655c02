/*
 * Batched block sets: N independent prepared problems (same block size,
 * MG level count and domain boundary spec) merged into one block set, so
 * that every smoother / apply / transfer launch covers the blocks of all
 * problems.  Block ids of problem p are shifted by the number of blocks of
 * problems 0..p-1; face neighbours come from an explicit table (-1 where a
 * problem's raster ends), so problems never see each other.  The MG0 pool
 * holds the PCG-active level-0 blocks of problem 0, then of problem 1, ...,
 * then every other block, so each problem's PCG vectors are one contiguous
 * range.  No reference counterpart: the reference solves one grid per call
 * (src/msbg3.cpp:3262-3271).
 */
#include <algorithm>
#include <cstring>

#include "msbgx_internal.h"

namespace msbgx {

int buildBatchGrid(Grid &m, const std::vector<const Grid *> &probs)
{
  if (probs.empty()) {
    setError("empty batch");
    return MSBGX_EINVAL;
  }
  const Grid &g0 = *probs[0];
  for (const Grid *p : probs) {
    if (!p->prepared || p->bsx0 != g0.bsx0 || p->nLevels != g0.nLevels ||
        memcmp(p->domainBc, g0.domainBc, sizeof(g0.domainBc)) != 0) {
      setError("batched problems must be prepared and share block size, level count and boundary spec");
      return MSBGX_EINVAL;
    }
  }
  m.bsx0 = g0.bsx0;
  m.nLevels = g0.nLevels;
  memcpy(m.domainBc, g0.domainBc, sizeof(m.domainBc));
  m.batched = true;
  m.nProb = (int)probs.size();
  m.sx0 = m.sy0 = m.sz0 = 0;
  m.nBlocks = 0;
  m.probBlock0.clear();
  m.probRaster.clear();
  for (const Grid *p : probs) {
    m.probBlock0.push_back(m.nBlocks);
    m.probRaster.push_back({p->nbx, p->nby, p->nbz});
    m.nBlocks += p->nBlocks;
  }
  /* no common raster: block ids run along x only; every neighbour lookup
   * goes through nbrTable */
  m.nbx = m.nBlocks;
  m.nby = m.nbz = 1;
  m.probOf.assign(m.nBlocks, 0);
  m.probOmega.clear();
  m.nbrTable.assign((size_t)m.nBlocks * 6, -1);
  m.map.clear();
  for (int q = 0; q < m.nProb; q++) {
    const Grid &p = *probs[q];
    const int s = m.probBlock0[q];
    m.probOmega.push_back(p.omega);
    for (int b = 0; b < p.nBlocks; b++) {
      m.probOf[s + b] = q;
      m.map.push_back(p.map[b]);
      int bx, by, bz;
      p.blockCoords(b, bx, by, bz);
      static const int o6[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
      for (int f = 0; f < 6; f++) {
        const int x = bx + o6[f][0], y = by + o6[f][1], z = bz + o6[f][2];
        m.nbrTable[(size_t)(s + b) * 6 + f] = p.inRaster(x, y, z) ? s + p.blockIndex(x, y, z) : -1;
      }
    }
  }
  m.omega = probs[0]->omega;
  m.prec = probs[0]->prec;

  for (int L = 0; L < m.nLevels; L++) {
    MgLevel &ml = m.mg[L];
    ml.level.clear();
    ml.bflags.clear();
    ml.relax.clear();
    for (int l = 0; l < kMaxLevels; l++) {
      ml.act[l].clear();
      ml.relaxCL[0][l].clear();
      ml.relaxCL[1][l].clear();
    }
    for (int q = 0; q < m.nProb; q++) {
      const MgLevel &pl = probs[q]->mg[L];
      const int s = m.probBlock0[q];
      ml.level.insert(ml.level.end(), pl.level.begin(), pl.level.end());
      ml.bflags.insert(ml.bflags.end(), pl.bflags.begin(), pl.bflags.end());
      for (int b : pl.relax) ml.relax.push_back(s + b);
      for (int l = 0; l < kMaxLevels; l++) {
        for (int b : pl.act[l]) ml.act[l].push_back(s + b);
        for (int c = 0; c < 2; c++)
          for (int b : pl.relaxCL[c][l]) ml.relaxCL[c][l].push_back(s + b);
      }
    }
    ml.anyRelax = !ml.relax.empty();
    /* slot order: L = 0 puts every problem's PCG-active blocks first */
    std::vector<int> order;
    order.reserve(m.nBlocks);
    if (L == 0) {
      m.active0.clear();
      m.probActive0.clear();
      for (int q = 0; q < m.nProb; q++) {
        m.probActive0.push_back((int)probs[q]->active0.size());
        for (int b : probs[q]->active0) order.push_back(m.probBlock0[q] + b);
      }
      m.active0 = order;
      m.nActive0 = (int)order.size();
      std::vector<char> seen(m.nBlocks, 0);
      for (int b : order) seen[b] = 1;
      for (int b = 0; b < m.nBlocks; b++)
        if (!seen[b]) order.push_back(b);
    } else {
      for (int b = 0; b < m.nBlocks; b++) order.push_back(b);
    }
    ml.off.assign(m.nBlocks, 0);
    int64_t o = 0;
    for (int b : order) {
      ml.off[b] = o;
      const int64_t e = m.bsxL(ml.level[b]);
      o += e * e * e;
    }
    ml.ncells = o;
    ml.flags.assign(ml.ncells, 0);
    for (int d = 0; d < 3; d++) ml.W[d].assign(ml.ncells, 0.0f);
    ml.D.assign(ml.ncells, 0.0f);
    for (int q = 0; q < m.nProb; q++) {
      const MgLevel &pl = probs[q]->mg[L];
      const int s = m.probBlock0[q];
      for (int b = 0; b < probs[q]->nBlocks; b++) {
        const int64_t e = m.bsxL(pl.level[b]), n = e * e * e, src = pl.off[b], dst = ml.off[s + b];
        std::copy(pl.flags.begin() + src, pl.flags.begin() + src + n, ml.flags.begin() + dst);
        for (int d = 0; d < 3; d++) std::copy(pl.W[d].begin() + src, pl.W[d].begin() + src + n, ml.W[d].begin() + dst);
        std::copy(pl.D.begin() + src, pl.D.begin() + src + n, ml.D.begin() + dst);
      }
    }
  }
  for (int L = 0; L < kMaxLevels; L++) m.aliasEstablished[L] = false;
  return MSBGX_OK;
}

}  // namespace msbgx
