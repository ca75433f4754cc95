/* CPU oracle of the Toucan cycle-simulation path (test infrastructure only;
 * see oracle.h). Plain C11 with GCC's unsigned __int128.
 *
 * Every evaluation routine cites the reference lines it restates. Where the
 * reference is undefined the oracle defines a value and counts the event:
 *   - LUT position >= 4642 (reads past __constant__ lutContent) reads 0;
 *   - memory address >= memPoolSize reads 0 / drops the write;
 *   - a value-pool slot read before it was written in the current visit
 *     (uninitialised shared memory, SimEval.cu:775-782) reads its last value
 *     and is counted.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define LUT_SIZE 4642u

typedef unsigned __int128 u128;

typedef struct { uint16_t lut, op0, op1, op2; } top_op;
typedef struct { uint8_t sid; uint16_t result; } last_op;
typedef struct {
  uint16_t vecBase, vecLength, offset; uint8_t isConst;
  uint16_t idx[4]; uint16_t outRange, result;
} vecread_op;
typedef struct { uint16_t v1, v2, len; uint8_t op, cflags; uint16_t result; } vecop_op;
typedef struct { uint8_t multi; uint64_t memBase; uint16_t en, addrVec, result; } memread_op;
typedef struct { uint8_t multi; uint64_t memBase; uint16_t addrVec, dat, en; } memwrite_op;

typedef struct {
  int isLut;
  uint32_t nTop; top_op *top;
  uint32_t nMid; uint32_t *midN; uint32_t **mid;
  uint32_t nLast; last_op *last;
  uint32_t nVr; vecread_op *vr;
  uint32_t nVo; vecop_op *vo;
  uint32_t nMr; memread_op *mr;
} mpart;

typedef struct {
  uint32_t vpSize;
  uint32_t nConst; uint8_t *constVec;
  uint32_t nRegRead; uint32_t *rrSrc; uint16_t *rrDst;
  uint32_t nExgRead; uint32_t *erSrc; uint16_t *erDst;
  uint32_t nLevels; uint32_t *levelN; mpart **levels;
  uint32_t rwReg; uint16_t rwDat, rwCount;
  uint32_t nMemWrite; memwrite_op *mw;
  uint32_t nPrint; uint16_t *prEn, *prMsg;
  uint32_t nStop; uint16_t *stEn;
  uint32_t ewExg; uint16_t ewDat, ewCount;
} part;

struct orc_prog {
  uint8_t lut[LUT_SIZE];
  uint32_t regPoolSize, memPoolSize, exchangePoolSize;
  uint32_t nParts; part *parts;
  uint32_t nRegion[2];
  uint32_t maxVp;
  uint32_t nMsgs;
};

struct orc_state {
  const orc_prog *p;
  uint8_t *reg, *mem, *exg;
  uint8_t *vp;        /* value pool of the partition being visited */
  uint32_t *written;  /* per slot: visit stamp of its last write */
  uint32_t visit;
  uint32_t *memEpoch; /* region epoch of the last write per memory nibble */
  uint32_t *memPart;  /* partition of the last access per memory nibble */
  uint32_t *memREpoch;
  uint64_t cycles; int stop; uint64_t stopCycle;
  uint32_t nIn; uint32_t *inIds; uint64_t xs;
  uint64_t viol[4];
  uint64_t nLog, capLog; uint64_t *logCycle; uint32_t *logMsg;
  /* signal probes: per partition a CSR list of value-pool slots whose value
   * at the end of the partition's visit is kept in probeVal */
  uint32_t nProbe; uint32_t *probeOff, *probeSlot, *probeIdx; uint8_t *probeVal;
};

/* ------------------------------------------------------------------ */
/* parser: src/Serializer.cpp:199-259 (design), :106-194 (parts)       */

typedef struct { const uint8_t *b; size_t n, pos; int bad; char *err; size_t errlen; } rd;

static int need(rd *r, size_t k, const char *what) {
  if (r->bad) return 0;
  if (k > r->n - r->pos) {
    r->bad = 1;
    snprintf(r->err, r->errlen, "truncated at byte %zu reading %s", r->pos, what);
    return 0;
  }
  return 1;
}
static uint64_t rdu(rd *r, int bytes, const char *what) {
  uint64_t v = 0;
  if (!need(r, (size_t)bytes, what)) return 0;
  memcpy(&v, r->b + r->pos, (size_t)bytes);
  r->pos += (size_t)bytes;
  return v;
}
static uint64_t rdcount(rd *r, size_t elem, const char *what) {
  uint64_t c = rdu(r, 8, what);
  if (r->bad) return 0;
  if (elem && c > (r->n - r->pos) / elem) {
    r->bad = 1;
    snprintf(r->err, r->errlen, "count %llu of %s exceeds input", (unsigned long long)c, what);
    return 0;
  }
  return c;
}
static const uint8_t *rdraw(rd *r, size_t k, const char *what) {
  if (!need(r, k, what)) return NULL;
  const uint8_t *p = r->b + r->pos;
  r->pos += k;
  return p;
}
static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s ? s : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}
static uint32_t up4(uint32_t x) { return (x + 3u) & ~3u; }
static uint16_t g16(const uint8_t *p) { uint16_t v; memcpy(&v, p, 2); return v; }
static uint32_t g32(const uint8_t *p) { uint32_t v; memcpy(&v, p, 4); return v; }
static uint64_t g64(const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }

static int load_mpart(rd *r, mpart *m) {
  m->isLut = (int)rdu(r, 1, "micropart kind") != 0;
  if (m->isLut) {
    uint64_t n = rdcount(r, 8, "top");
    const uint8_t *q = rdraw(r, n * 8, "top");
    if (r->bad) return 0;
    m->nTop = (uint32_t)n; m->top = xcalloc(n, sizeof(top_op));
    for (uint64_t i = 0; i < n; ++i) {
      m->top[i].lut = g16(q + 8 * i); m->top[i].op0 = g16(q + 8 * i + 2);
      m->top[i].op1 = g16(q + 8 * i + 4); m->top[i].op2 = g16(q + 8 * i + 6);
    }
    uint64_t nm = rdcount(r, 8, "middle levels");
    if (r->bad) return 0;
    m->nMid = (uint32_t)nm; m->midN = xcalloc(nm, 4); m->mid = xcalloc(nm, sizeof(uint32_t *));
    for (uint64_t l = 0; l < nm; ++l) {
      uint64_t k = rdcount(r, 4, "middle level");
      const uint8_t *qq = rdraw(r, k * 4, "middle level");
      if (r->bad) return 0;
      m->midN[l] = (uint32_t)k; m->mid[l] = xcalloc(k, 4);
      for (uint64_t i = 0; i < k; ++i) m->mid[l][i] = g32(qq + 4 * i);
    }
    n = rdcount(r, 4, "last");
    q = rdraw(r, n * 4, "last");
    if (r->bad) return 0;
    m->nLast = (uint32_t)n; m->last = xcalloc(n, sizeof(last_op));
    for (uint64_t i = 0; i < n; ++i) { m->last[i].sid = q[4 * i]; m->last[i].result = g16(q + 4 * i + 2); }
  } else {
    uint64_t n = rdcount(r, 20, "vecRead");
    const uint8_t *q = rdraw(r, n * 20, "vecRead");
    if (r->bad) return 0;
    m->nVr = (uint32_t)n; m->vr = xcalloc(n, sizeof(vecread_op));
    for (uint64_t i = 0; i < n; ++i) {
      const uint8_t *e = q + 20 * i;
      vecread_op *o = &m->vr[i];
      o->vecBase = g16(e); o->vecLength = g16(e + 2); o->offset = g16(e + 4); o->isConst = e[6];
      for (int k = 0; k < 4; ++k) o->idx[k] = g16(e + 8 + 2 * k);
      o->outRange = g16(e + 16); o->result = g16(e + 18);
    }
    n = rdcount(r, 10, "vecOp");
    q = rdraw(r, n * 10, "vecOp");
    if (r->bad) return 0;
    m->nVo = (uint32_t)n; m->vo = xcalloc(n, sizeof(vecop_op));
    for (uint64_t i = 0; i < n; ++i) {
      const uint8_t *e = q + 10 * i;
      vecop_op *o = &m->vo[i];
      o->v1 = g16(e); o->v2 = g16(e + 2); o->len = g16(e + 4); o->op = e[6]; o->cflags = e[7]; o->result = g16(e + 8);
    }
    n = rdcount(r, 24, "memRead");
    q = rdraw(r, n * 24, "memRead");
    if (r->bad) return 0;
    m->nMr = (uint32_t)n; m->mr = xcalloc(n, sizeof(memread_op));
    for (uint64_t i = 0; i < n; ++i) {
      const uint8_t *e = q + 24 * i;
      memread_op *o = &m->mr[i];
      o->multi = e[0]; o->memBase = g64(e + 8); o->en = g16(e + 16); o->addrVec = g16(e + 18); o->result = g16(e + 20);
    }
  }
  return !r->bad;
}

static int load_part(rd *r, part *p) {
  uint64_t n = rdcount(r, 1, "valuePool");
  rdraw(r, n, "valuePool");            /* host consts, never loaded on the device */
  p->vpSize = (uint32_t)rdu(r, 4, "valuePoolSize");
  n = rdcount(r, 1, "constVecPool");
  const uint8_t *q = rdraw(r, n, "constVecPool");
  if (r->bad) return 0;
  p->nConst = (uint32_t)n; p->constVec = xcalloc(n, 1);
  if (n) memcpy(p->constVec, q, n);
  n = rdcount(r, 8, "regRead");
  q = rdraw(r, n * 8, "regRead");
  if (r->bad) return 0;
  p->nRegRead = (uint32_t)n; p->rrSrc = xcalloc(n, 4); p->rrDst = xcalloc(n, 2);
  for (uint64_t i = 0; i < n; ++i) { p->rrSrc[i] = g32(q + 8 * i); p->rrDst[i] = g16(q + 8 * i + 4); }
  n = rdcount(r, 8, "exchangeRead");
  q = rdraw(r, n * 8, "exchangeRead");
  if (r->bad) return 0;
  p->nExgRead = (uint32_t)n; p->erSrc = xcalloc(n, 4); p->erDst = xcalloc(n, 2);
  for (uint64_t i = 0; i < n; ++i) { p->erSrc[i] = g32(q + 8 * i); p->erDst[i] = g16(q + 8 * i + 4); }
  n = rdcount(r, 8, "exec levels");
  if (r->bad) return 0;
  p->nLevels = (uint32_t)n; p->levelN = xcalloc(n, 4); p->levels = xcalloc(n, sizeof(mpart *));
  for (uint64_t l = 0; l < n; ++l) {
    uint64_t k = rdcount(r, 1, "microparts");
    if (r->bad) return 0;
    p->levelN[l] = (uint32_t)k; p->levels[l] = xcalloc(k, sizeof(mpart));
    for (uint64_t i = 0; i < k; ++i)
      if (!load_mpart(r, &p->levels[l][i])) return 0;
  }
  q = rdraw(r, 8, "regWrite");
  if (r->bad) return 0;
  p->rwReg = g32(q); p->rwDat = g16(q + 4); p->rwCount = g16(q + 6);
  n = rdcount(r, 24, "memWrite");
  q = rdraw(r, n * 24, "memWrite");
  if (r->bad) return 0;
  p->nMemWrite = (uint32_t)n; p->mw = xcalloc(n, sizeof(memwrite_op));
  for (uint64_t i = 0; i < n; ++i) {
    const uint8_t *e = q + 24 * i;
    p->mw[i].multi = e[0]; p->mw[i].memBase = g64(e + 8);
    p->mw[i].addrVec = g16(e + 16); p->mw[i].dat = g16(e + 18); p->mw[i].en = g16(e + 20);
  }
  n = rdcount(r, 4, "print");
  q = rdraw(r, n * 4, "print");
  if (r->bad) return 0;
  p->nPrint = (uint32_t)n; p->prEn = xcalloc(n, 2); p->prMsg = xcalloc(n, 2);
  for (uint64_t i = 0; i < n; ++i) { p->prEn[i] = g16(q + 4 * i); p->prMsg[i] = g16(q + 4 * i + 2); }
  n = rdcount(r, 2, "stop");
  q = rdraw(r, n * 2, "stop");
  if (r->bad) return 0;
  p->nStop = (uint32_t)n; p->stEn = xcalloc(n, 2);
  for (uint64_t i = 0; i < n; ++i) p->stEn[i] = g16(q + 2 * i);
  q = rdraw(r, 8, "exchangeWrite");
  if (r->bad) return 0;
  p->ewExg = g32(q); p->ewDat = g16(q + 4); p->ewCount = g16(q + 6);
  return 1;
}

static int fail(char *err, size_t errlen, const char *msg) {
  snprintf(err, errlen, "%s", msg);
  return -1;
}

int orc_prog_load(const uint8_t *buf, size_t size, orc_prog **out, char *err, size_t errlen) {
  char dummy[8];
  if (!err) { err = dummy; errlen = sizeof dummy; }
  rd r = {buf, size, 0, 0, err, errlen};
  orc_prog *p = xcalloc(1, sizeof(orc_prog));
  p->regPoolSize = (uint32_t)rdu(&r, 4, "regPoolSize");
  p->memPoolSize = (uint32_t)rdu(&r, 4, "memPoolSize");
  p->exchangePoolSize = (uint32_t)rdu(&r, 4, "exchangePoolSize");
  rdu(&r, 1, "shouldStop");
  uint64_t n = rdcount(&r, 1, "lut");
  const uint8_t *q = rdraw(&r, n, "lut");
  if (!r.bad && n != LUT_SIZE) { free(p); return fail(err, errlen, "LUT size is not 4642 (SimEval.cu:1122)"); }
  if (q) memcpy(p->lut, q, LUT_SIZE);
  n = rdcount(&r, 1, "regPool"); rdraw(&r, n, "regPool");
  n = rdcount(&r, 1, "memPool"); rdraw(&r, n, "memPool");
  n = rdcount(&r, 8, "parts");
  if (r.bad) { free(p); return -1; }
  p->nParts = (uint32_t)n; p->parts = xcalloc(n, sizeof(part));
  for (uint64_t i = 0; i < n; ++i)
    if (!load_part(&r, &p->parts[i])) { *out = NULL; return -1; }
  n = rdcount(&r, 8, "regions");
  if (r.bad) return -1;
  if (n != 2) return fail(err, errlen, "exactly 2 regions are supported (SimEval.cu:1292-1295)");
  uint32_t expect = 0;
  for (int g = 0; g < 2; ++g) {
    uint64_t k = rdcount(&r, 4, "region");
    q = rdraw(&r, k * 4, "region");
    if (r.bad) return -1;
    for (uint64_t i = 0; i < k; ++i)
      if (g32(q + 4 * i) != expect++) return fail(err, errlen, "region partition ids must be contiguous from 0 (SimEval.cu:1298-1302)");
    p->nRegion[g] = (uint32_t)k;
  }
  if (expect != p->nParts) return fail(err, errlen, "regions do not cover every partition");
  n = rdcount(&r, 8, "messages");
  for (uint64_t i = 0; i < n && !r.bad; ++i) { uint64_t k = rdcount(&r, 1, "message"); rdraw(&r, k, "message"); }
  if (r.bad) return -1;
  p->nMsgs = (uint32_t)n;
  for (uint32_t i = 0; i < p->nParts; ++i) {
    part *pp = &p->parts[i];
    if (pp->vpSize > 65535) return fail(err, errlen, "value pool over 65535 slots (ToucanSim.cu:153)");
    if (pp->vpSize > p->maxVp) p->maxVp = pp->vpSize;
    if ((pp->nRegRead == 0) == (pp->nExgRead == 0))
      return fail(err, errlen, "a partition needs exactly one of regRead/exchangeRead (SimEval.cu:1183)");
    for (uint32_t k = 0; k < pp->nRegRead; ++k)
      if (pp->rrSrc[k] >= p->regPoolSize) return fail(err, errlen, "regRead source out of range");
    for (uint32_t k = 0; k < pp->nExgRead; ++k)
      if (pp->erSrc[k] >= p->exchangePoolSize) return fail(err, errlen, "exchangeRead source out of range");
    if (pp->ewCount) {
      uint32_t w = ((uint32_t)pp->ewCount + 3) / 4;
      if (((pp->ewExg >> 2) + w) * 4 > up4(p->exchangePoolSize) || ((pp->ewDat >> 2) + w) * 4 > 65536)
        return fail(err, errlen, "exchange write out of range");
    }
    if (pp->rwCount) {
      uint32_t w = ((uint32_t)pp->rwCount + 3) / 4;
      if (((pp->rwReg >> 2) + w) * 4 > up4(p->regPoolSize) || ((pp->rwDat >> 2) + w) * 4 > 65536)
        return fail(err, errlen, "register write out of range");
    }
    for (uint32_t k = 0; k < pp->nPrint; ++k)
      if (pp->prMsg[k] >= p->nMsgs) return fail(err, errlen, "print message index out of range");
  }
  *out = p;
  return 0;
}

static void free_mpart(mpart *m) {
  free(m->top);
  for (uint32_t l = 0; l < m->nMid; ++l) free(m->mid[l]);
  free(m->mid); free(m->midN); free(m->last); free(m->vr); free(m->vo); free(m->mr);
}

void orc_prog_free(orc_prog *p) {
  if (!p) return;
  for (uint32_t i = 0; i < p->nParts; ++i) {
    part *pp = &p->parts[i];
    for (uint32_t l = 0; l < pp->nLevels; ++l) {
      for (uint32_t k = 0; k < pp->levelN[l]; ++k) free_mpart(&pp->levels[l][k]);
      free(pp->levels[l]);
    }
    free(pp->levels); free(pp->levelN); free(pp->constVec); free(pp->rrSrc); free(pp->rrDst);
    free(pp->erSrc); free(pp->erDst); free(pp->mw); free(pp->prEn); free(pp->prMsg); free(pp->stEn);
  }
  free(p->parts);
  free(p);
}

uint32_t orc_pool_size(const orc_prog *p, int which) {
  return which == 0 ? p->regPoolSize : which == 1 ? p->memPoolSize : p->exchangePoolSize;
}
uint32_t orc_num_parts(const orc_prog *p, int region) { return p->nRegion[region & 1]; }

/* ------------------------------------------------------------------ */
/* state                                                               */

int orc_state_new(const orc_prog *p, orc_state **out) {
  orc_state *s = xcalloc(1, sizeof(orc_state));
  s->p = p;
  /* pools are word-copied (SimEval.cu:658-665,751-758): round up to 4 */
  s->reg = xcalloc(up4(p->regPoolSize) + 4, 1);
  s->mem = xcalloc((size_t)p->memPoolSize + 1, 1);
  s->exg = xcalloc(up4(p->exchangePoolSize) + 4, 1);
  s->vp = xcalloc(65536 + 64, 1);
  s->written = xcalloc(65536 + 64, 4);
  s->memEpoch = xcalloc((size_t)p->memPoolSize + 1, 4);
  s->memPart = xcalloc((size_t)p->memPoolSize + 1, 4);
  s->memREpoch = xcalloc((size_t)p->memPoolSize + 1, 4);
  s->stopCycle = UINT64_MAX;
  *out = s;
  return 0;
}

void orc_state_free(orc_state *s) {
  if (!s) return;
  free(s->reg); free(s->mem); free(s->exg); free(s->vp); free(s->written);
  free(s->memEpoch); free(s->memPart); free(s->memREpoch); free(s->inIds);
  free(s->logCycle); free(s->logMsg);
  free(s->probeOff); free(s->probeSlot); free(s->probeIdx); free(s->probeVal);
  free(s);
}

int orc_state_copy_pool(orc_state *s, int which, int to_state, uint8_t *buf, size_t n) {
  uint8_t *pool = which == 0 ? s->reg : which == 1 ? s->mem : s->exg;
  size_t sz = orc_pool_size(s->p, which);
  if (n > sz) return -1;
  if (to_state) memcpy(pool, buf, n); else memcpy(buf, pool, n);
  return 0;
}

/* xorshift64, Marsaglia (2003) "Xorshift RNGs", shifts (13, 7, 17). */
uint64_t orc_xorshift64(uint64_t x) {
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  return x;
}

/* SplitMix64 (Steele, Lea, Flood 2014; constants as published by Vigna). */
uint64_t orc_splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int orc_set_stimulus(orc_state *s, const uint32_t *ids, uint32_t n, uint64_t seed, uint32_t instance) {
  for (uint32_t i = 0; i < n; ++i)
    if (ids[i] >= s->p->regPoolSize) return -1;
  free(s->inIds);
  s->inIds = NULL;
  s->nIn = n;
  if (n) {
    s->inIds = xcalloc(n, 4);
    memcpy(s->inIds, ids, (size_t)n * 4);
  }
  s->xs = orc_splitmix64(seed + instance);
  if (s->xs == 0) s->xs = 0x9E3779B97F4A7C15ull;
  return 0;
}

uint64_t orc_cycles(const orc_state *s) { return s->cycles; }
int orc_stopped(const orc_state *s) { return s->stop; }
uint64_t orc_stop_cycle(const orc_state *s) { return s->stopCycle; }
void orc_violations(const orc_state *s, uint64_t out[4]) { memcpy(out, s->viol, sizeof(s->viol)); }

uint64_t orc_print_log(const orc_state *s, uint64_t *cycles, uint32_t *msgs, uint64_t cap) {
  for (uint64_t i = 0; i < s->nLog && i < cap; ++i) { cycles[i] = s->logCycle[i]; msgs[i] = s->logMsg[i]; }
  return s->nLog;
}

int orc_set_probes(orc_state *s, const uint32_t *parts, const uint32_t *slots, uint32_t n) {
  const orc_prog *p = s->p;
  for (uint32_t i = 0; i < n; ++i)
    if (parts[i] >= p->nParts || slots[i] >= 65536) return -1;
  free(s->probeOff); free(s->probeSlot); free(s->probeIdx); free(s->probeVal);
  s->probeOff = NULL; s->probeSlot = s->probeIdx = NULL; s->probeVal = NULL;
  s->nProbe = n;
  if (!n) return 0;
  s->probeOff = xcalloc(p->nParts + 1, 4);
  s->probeSlot = xcalloc(n, 4);
  s->probeIdx = xcalloc(n, 4);
  s->probeVal = xcalloc(n, 1);
  for (uint32_t i = 0; i < n; ++i) s->probeOff[parts[i] + 1]++;
  for (uint32_t q = 0; q < p->nParts; ++q) s->probeOff[q + 1] += s->probeOff[q];
  uint32_t *fill = xcalloc(p->nParts, 4);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t at = s->probeOff[parts[i]] + fill[parts[i]]++;
    s->probeSlot[at] = slots[i];
    s->probeIdx[at] = i;
  }
  free(fill);
  return 0;
}

void orc_read_probes(const orc_state *s, uint8_t *out) {
  if (s->nProbe) memcpy(out, s->probeVal, s->nProbe);
}

/* Slots every visit of partition `part` writes, whatever the state: L0
 * destinations and the results of LUT, VecRead and VecOp microparts
 * (a MemRead writes only when enabled, so its results are left out). */
uint32_t orc_written_slots(const orc_prog *p, uint32_t pi, uint32_t *out, uint32_t cap) {
  if (pi >= p->nParts) return 0;
  const part *q = &p->parts[pi];
  uint8_t *seen = xcalloc(65536, 1);
  uint32_t n = 0;
#define PUT(x) do { uint32_t v_ = (x) & 0xFFFFu; if (!seen[v_]) { seen[v_] = 1; if (n < cap) out[n] = v_; n++; } } while (0)
  for (uint32_t i = 0; i < q->nRegRead; ++i) PUT(q->rrDst[i]);
  for (uint32_t i = 0; i < q->nExgRead; ++i) PUT(q->erDst[i]);
  for (uint32_t l = 0; l < q->nLevels; ++l)
    for (uint32_t k = 0; k < q->levelN[l]; ++k) {
      const mpart *m = &q->levels[l][k];
      if (m->isLut) {
        for (uint32_t i = 0; i < m->nLast && i < 32; ++i) PUT(m->last[i].result);
      } else if (m->nVr) {
        for (uint32_t i = 0; i < m->nVr && i < 32; ++i) PUT(m->vr[i].result);
      } else if (m->nVo) {
        for (uint32_t i = 0; i < m->nVo && i < 32; ++i) {
          const vecop_op *o = &m->vo[i];
          if (o->op <= 2)
            for (uint32_t j = 0; j < o->len; ++j) PUT((uint16_t)(o->result + j));
          else
            PUT(o->result);
        }
      }
    }
#undef PUT
  free(seen);
  return n;
}

/* ------------------------------------------------------------------ */
/* evaluation                                                          */

static inline uint8_t vp_get(orc_state *s, uint32_t slot) {
  if (s->written[slot] != s->visit) s->viol[0]++;
  return s->vp[slot];
}
static inline void vp_put(orc_state *s, uint32_t slot, uint8_t v) {
  s->vp[slot] = v;
  s->written[slot] = s->visit;
}
/* operand rule shared by top-level ops, VecRead indices, enables
 * (SimEval.cu:345-347,458-461,476,559,674,677): < 16 is an immediate */
static inline uint32_t imm16(orc_state *s, uint16_t v) { return v < 16 ? v : vp_get(s, v); }

/* LUT lookup (SimEval.cu:349-356,406-413): index 0 passes op2 through, else
 * lutContent[(u16)(lut + (op0<<8 | op1<<4 | op2))]. */
static inline uint8_t lut_eval(orc_state *s, uint32_t lut, uint32_t a, uint32_t b, uint32_t c) {
  if (lut == 0) return (uint8_t)c;
  uint16_t pos = (uint16_t)(lut + (((uint32_t)(uint16_t)a << 8) | (b << 4) | c));
  if (pos >= LUT_SIZE) { s->viol[3]++; return 0; }
  return s->p->lut[pos];
}

/* Memory address of SimEval.cu:561-570 / 680-690. */
static inline uint64_t mem_addr(orc_state *s, uint16_t addrVec, uint8_t multi, uint64_t base) {
  uint32_t addr = 0;
  for (uint32_t j = 0; j < 8; ++j) addr |= (uint32_t)vp_get(s, (uint32_t)addrVec + j) << (4 * j);
  if (multi) addr <<= 2;
  return base + addr;
}

/* sign_extend_128 (SimEval.cu:222-228); lengths 0 and 32 need no extension */
static u128 sext(u128 v, uint32_t len) {
  if (len == 0 || len >= 32) return v;
  uint32_t bit = len * 4 - 1;
  if (!((v >> bit) & 1)) return v;
  u128 mask = (((u128)1) << (bit + 1)) - 1;
  return v | ~mask;
}

static void eval_lut_mpart(orc_state *s, const mpart *m) {
  /* SimEval.cu:319-448: one 32-lane warp; lanes past a level's size hold 0 */
  uint8_t prev[32], cur[32];
  memset(prev, 0, sizeof prev);
  for (uint32_t i = 0; i < m->nTop && i < 32; ++i) {
    const top_op *o = &m->top[i];
    uint32_t c = imm16(s, o->op2);
    uint32_t a = 0, b = 0;
    if (o->lut != 0) { a = imm16(s, o->op0); b = imm16(s, o->op1); }
    prev[i] = lut_eval(s, o->lut, a, b, c);
  }
  for (uint32_t l = 0; l < m->nMid; ++l) {
    for (uint32_t i = 0; i < 32; ++i) {
      uint32_t w = i < m->midN[l] ? m->mid[l][i] : 0;
      uint32_t o2 = w & 63, o1 = (w >> 6) & 63, o0 = (w >> 12) & 63, lut = w >> 18;
      /* shuffle source lane is taken mod 32 (CUDA __shfl_sync width rule) */
      uint32_t a = o0 < 32 ? o0 : prev[o0 & 31];
      uint32_t b = o1 < 32 ? o1 : prev[o1 & 31];
      uint32_t c = o2 < 32 ? o2 : prev[o2 & 31];
      cur[i] = lut_eval(s, lut, a, b, c);
    }
    memcpy(prev, cur, sizeof prev);
  }
  /* reads of the last level all see the final middle level, then write */
  uint8_t vals[32];
  for (uint32_t i = 0; i < m->nLast && i < 32; ++i) {
    uint32_t sid = m->last[i].sid;
    vals[i] = sid < 32 ? (uint8_t)sid : prev[sid & 31];
  }
  for (uint32_t i = 0; i < m->nLast && i < 32; ++i) vp_put(s, m->last[i].result, vals[i]);
}

static void eval_vecread(orc_state *s, const part *pp, const mpart *m) {
  /* SimEval.cu:449-481; all lanes read before any lane writes */
  uint8_t vals[32];
  uint32_t n = m->nVr < 32 ? m->nVr : 32;
  for (uint32_t i = 0; i < n; ++i) {
    const vecread_op *o = &m->vr[i];
    uint32_t i0 = imm16(s, o->idx[0]), i1 = imm16(s, o->idx[1]), i2 = imm16(s, o->idx[2]), i3 = imm16(s, o->idx[3]);
    uint32_t off = ((i0 << 12) | (i1 << 8) | (i2 << 4) | i3) + o->offset;
    if (off < o->vecLength) {
      uint32_t at = (uint32_t)o->vecBase + off;
      if (o->isConst) vals[i] = at < pp->nConst ? pp->constVec[at] : 0;
      else vals[i] = at < 65536 ? vp_get(s, at) : 0;
    } else {
      vals[i] = (uint8_t)imm16(s, o->outRange);
    }
  }
  for (uint32_t i = 0; i < n; ++i) vp_put(s, m->vr[i].result, vals[i]);
}

static void eval_vecop(orc_state *s, const part *pp, const mpart *m) {
  /* SimEval.cu:482-548 */
  uint32_t n = m->nVo < 32 ? m->nVo : 32;
  u128 res[32];
  uint8_t scalar[32];
  for (uint32_t i = 0; i < n; ++i) {
    const vecop_op *o = &m->vo[i];
    int c1 = (o->cflags & 2) != 0, c2 = (o->cflags & 1) != 0;
    u128 v1 = 0, v2 = 0;
    for (uint32_t k = 0; k < o->len && k < 32; ++k) {
      uint32_t a1 = (uint32_t)o->v1 + k, a2 = (uint32_t)o->v2 + k;
      u128 s1 = c1 ? (a1 < pp->nConst ? pp->constVec[a1] : 0) : vp_get(s, a1);
      u128 s2 = c2 ? (a2 < pp->nConst ? pp->constVec[a2] : 0) : vp_get(s, a2);
      v1 |= s1 << (4 * k);
      v2 |= s2 << (4 * k);
    }
    v1 = sext(v1, o->len);
    v2 = sext(v2, o->len);
    __int128 sv1 = (__int128)v1, sv2 = (__int128)v2;
    res[i] = v1;
    scalar[i] = 0;
    switch (o->op) {
      case 0: res[i] = v1 + v2; break;
      case 1: res[i] = v1 - v2; break;
      case 2: res[i] = v1 * v2; break;
      case 3: scalar[i] = sv1 == sv2; break;
      case 4: scalar[i] = sv1 < sv2; break;  /* signed, after sign extension */
      case 5: scalar[i] = sv1 <= sv2; break;
      default: break;
    }
  }
  for (uint32_t i = 0; i < n; ++i) {
    const vecop_op *o = &m->vo[i];
    if (o->op <= 2) {
      u128 r = res[i];
      for (uint32_t k = 0; k < o->len; ++k) {
        vp_put(s, (uint16_t)(o->result + k), (uint8_t)(r & 15));
        r >>= 4;
      }
    } else {
      vp_put(s, o->result, scalar[i]);  /* unknown op names write 0 */
    }
  }
}

static void eval_memread(orc_state *s, uint32_t partId, uint32_t epoch, const mpart *m) {
  /* SimEval.cu:550-575: a disabled read writes nothing */
  uint32_t n = m->nMr < 32 ? m->nMr : 32;
  uint8_t vals[32];
  int en[32];
  for (uint32_t i = 0; i < n; ++i) {
    const memread_op *o = &m->mr[i];
    en[i] = imm16(s, o->en) != 0;
    if (!en[i]) continue;
    uint64_t at = mem_addr(s, o->addrVec, o->multi, o->memBase);
    if (at >= s->p->memPoolSize) { s->viol[2]++; vals[i] = 0; continue; }
    if (s->memEpoch[at] == epoch) s->viol[1]++;
    s->memREpoch[at] = epoch;
    s->memPart[at] = partId;
    vals[i] = s->mem[at];
  }
  for (uint32_t i = 0; i < n; ++i)
    if (en[i]) vp_put(s, m->mr[i].result, vals[i]);
}

static void word_copy(orc_state *s, uint8_t *pool, uint32_t dst, uint16_t dat, uint16_t count) {
  /* SimEval.cu:658-665,751-758: ceil(count/4) 4-nibble words, both ends aligned down */
  uint32_t words = ((uint32_t)count + 3) / 4;
  for (uint32_t i = 0; i < words; ++i)
    for (uint32_t k = 0; k < 4; ++k)
      pool[((dst >> 2) + i) * 4 + k] = vp_get(s, (((uint32_t)dat >> 2) + i) * 4 + k);
}

static void eval_partition(orc_state *s, uint32_t partId, uint32_t epoch, int enablePrint) {
  const part *pp = &s->p->parts[partId];
  if (++s->visit == 0) { memset(s->written, 0, (65536 + 64) * 4); s->visit = 1; }
  /* L0 (SimEval.cu:582-622) */
  for (uint32_t i = 0; i < pp->nRegRead; ++i) vp_put(s, pp->rrDst[i], s->reg[pp->rrSrc[i]]);
  for (uint32_t i = 0; i < pp->nExgRead; ++i) vp_put(s, pp->erDst[i], s->exg[pp->erSrc[i]]);
  /* exec levels in order, microparts in listed order (SimEval.cu:806-844) */
  for (uint32_t l = 0; l < pp->nLevels; ++l) {
    for (uint32_t k = 0; k < pp->levelN[l]; ++k) {
      const mpart *m = &pp->levels[l][k];
      if (m->isLut) eval_lut_mpart(s, m);
      else if (m->nVr) eval_vecread(s, pp, m);
      else if (m->nVo) eval_vecop(s, pp, m);
      else eval_memread(s, partId, epoch, m);
    }
  }
  /* signal probes: the slot values this visit leaves behind */
  if (s->probeOff)
    for (uint32_t k = s->probeOff[partId]; k < s->probeOff[partId + 1]; ++k)
      s->probeVal[s->probeIdx[k]] = s->vp[s->probeSlot[k]];
  /* last level (SimEval.cu:846-859) */
  if (pp->ewCount != 0) {
    word_copy(s, s->exg, pp->ewExg, pp->ewDat, pp->ewCount);
    return;
  }
  if (pp->rwCount != 0) word_copy(s, s->reg, pp->rwReg, pp->rwDat, pp->rwCount);
  for (uint32_t i = 0; i < pp->nMemWrite; ++i) {  /* SimEval.cu:670-693 */
    const memwrite_op *o = &pp->mw[i];
    if (imm16(s, o->en) == 0) continue;
    uint8_t dat = (uint8_t)imm16(s, o->dat);
    uint64_t at = mem_addr(s, o->addrVec, o->multi, o->memBase);
    if (at >= s->p->memPoolSize) { s->viol[2]++; continue; }
    if ((s->memEpoch[at] == epoch && s->memPart[at] != partId) ||
        (s->memREpoch[at] == epoch && s->memPart[at] != partId))
      s->viol[1]++;
    s->memEpoch[at] = epoch;
    s->memPart[at] = partId;
    s->mem[at] = dat;
  }
  for (uint32_t i = 0; i < pp->nPrint; ++i) {  /* SimEval.cu:696-704 */
    if (!enablePrint) continue;
    if (imm16(s, pp->prEn[i]) == 0) continue;
    if (s->nLog == s->capLog) {
      s->capLog = s->capLog ? 2 * s->capLog : 1024;
      s->logCycle = realloc(s->logCycle, s->capLog * 8);
      s->logMsg = realloc(s->logMsg, s->capLog * 4);
      if (!s->logCycle || !s->logMsg) abort();
    }
    s->logCycle[s->nLog] = s->cycles;
    s->logMsg[s->nLog] = pp->prMsg[i];
    s->nLog++;
  }
  for (uint32_t i = 0; i < pp->nStop; ++i) {  /* SimEval.cu:706-722: sticky */
    if (imm16(s, pp->stEn[i]) != 0) {
      if (!s->stop) s->stopCycle = s->cycles;
      s->stop = 1;
    }
  }
}

static void apply_stimulus(orc_state *s) {
  for (uint32_t d = 0; d * 64 < s->nIn; ++d) {
    s->xs = orc_xorshift64(s->xs);
    for (uint32_t j = 0; j < 64 && d * 64 + j < s->nIn; ++j)
      s->reg[s->inIds[d * 64 + j]] = (uint8_t)((s->xs >> j) & 1);
  }
}

int orc_run(orc_state *s, uint32_t cycles, int enable_print) {
  const orc_prog *p = s->p;
  for (uint32_t c = 0; c < cycles; ++c) {
    /* stimulus before the cycle */
    apply_stimulus(s);
    /* evalFreeRunningNCycles (SimEval.cu:889-995): region 0, barrier, region 1, barrier */
    uint32_t e0 = (uint32_t)(2 * (s->cycles & 0x7FFFFFFF) + 1);
    for (uint32_t i = 0; i < p->nRegion[0]; ++i) eval_partition(s, i, e0, enable_print);
    for (uint32_t i = p->nRegion[0]; i < p->nParts; ++i) eval_partition(s, i, e0 + 1, enable_print);
    s->cycles++;
  }
  return s->stop;
}

/* ------------------------------------------------------------------ */
/* multi-threaded run: the partitions of a region are taken one at a time */
/* from a shared counter by `threads` threads, with a barrier between    */
/* regions as the reference's grid sync (SimEval.cu:889-995). Each      */
/* thread has its own value pool and undefined-behaviour counters; the  */
/* pools are shared. Partitions of one region write disjoint pool words, */
/* so the pools, stop flag and stop cycle equal orc_run's; print entries */
/* of one cycle may come in another order. The cross-partition memory-  */
/* sharing counter (viol[1]) is only exact single-threaded.             */

typedef struct {
  orc_state w;         /* shallow copy of the main state with private scratch */
  orc_state *main;
  uint32_t tid, nThreads, cycles;
  int enablePrint;
  pthread_barrier_t *bar;
  uint32_t *next;      /* shared: next partition of region 0 / region 1 */
} orc_worker;

static void merge_cycle(orc_worker *ws, uint32_t n) {
  orc_state *m = ws[0].main;
  for (uint32_t t = 0; t < n; ++t) {
    orc_state *w = &ws[t].w;
    if (w->stop && !m->stop) { m->stop = 1; m->stopCycle = m->cycles; }
    w->stop = 0;
    for (uint64_t i = 0; i < w->nLog; ++i) {
      if (m->nLog == m->capLog) {
        m->capLog = m->capLog ? 2 * m->capLog : 1024;
        m->logCycle = realloc(m->logCycle, m->capLog * 8);
        m->logMsg = realloc(m->logMsg, m->capLog * 4);
        if (!m->logCycle || !m->logMsg) abort();
      }
      m->logCycle[m->nLog] = w->logCycle[i];
      m->logMsg[m->nLog] = w->logMsg[i];
      m->nLog++;
    }
    w->nLog = 0;
  }
  m->cycles++;
  ws[0].next[0] = 0;
  ws[0].next[1] = m->p->nRegion[0];
}

static void *worker_main(void *arg) {
  orc_worker *me = (orc_worker *)arg;
  orc_worker *all = me - me->tid;
  orc_state *m = me->main;
  const orc_prog *p = m->p;
  for (uint32_t c = 0; c < me->cycles; ++c) {
    if (me->tid == 0) apply_stimulus(m);
    pthread_barrier_wait(me->bar);
    me->w.cycles = m->cycles;
    uint32_t e0 = (uint32_t)(2 * (m->cycles & 0x7FFFFFFF) + 1);
    for (;;) {
      uint32_t i = __atomic_fetch_add(&me->next[0], 1, __ATOMIC_RELAXED);
      if (i >= p->nRegion[0]) break;
      eval_partition(&me->w, i, e0, me->enablePrint);
    }
    pthread_barrier_wait(me->bar);
    for (;;) {
      uint32_t i = __atomic_fetch_add(&me->next[1], 1, __ATOMIC_RELAXED);
      if (i >= p->nParts) break;
      eval_partition(&me->w, i, e0 + 1, me->enablePrint);
    }
    pthread_barrier_wait(me->bar);
    if (me->tid == 0) merge_cycle(all, me->nThreads);
  }
  return NULL;
}

int orc_run_threads(orc_state *s, uint32_t cycles, int enable_print, int threads) {
  if (threads <= 1) return orc_run(s, cycles, enable_print);
  uint32_t n = (uint32_t)threads;
  orc_worker *ws = xcalloc(n, sizeof(orc_worker));
  pthread_t *th = xcalloc(n, sizeof(pthread_t));
  pthread_barrier_t bar;
  pthread_barrier_init(&bar, NULL, n);
  uint32_t next[2] = {0, s->p->nRegion[0]};
  for (uint32_t t = 0; t < n; ++t) {
    ws[t].w = *s;
    ws[t].w.vp = xcalloc(65536 + 64, 1);
    ws[t].w.written = xcalloc(65536 + 64, 4);
    ws[t].w.visit = 0;
    memset(ws[t].w.viol, 0, sizeof(ws[t].w.viol));
    ws[t].w.nLog = ws[t].w.capLog = 0;
    ws[t].w.logCycle = NULL;
    ws[t].w.logMsg = NULL;
    ws[t].w.stop = 0;
    ws[t].main = s;
    ws[t].tid = t;
    ws[t].nThreads = n;
    ws[t].cycles = cycles;
    ws[t].enablePrint = enable_print;
    ws[t].bar = &bar;
    ws[t].next = next;
  }
  for (uint32_t t = 1; t < n; ++t)
    if (pthread_create(&th[t], NULL, worker_main, &ws[t]) != 0) abort();
  worker_main(&ws[0]);
  for (uint32_t t = 1; t < n; ++t) pthread_join(th[t], NULL);
  for (uint32_t t = 0; t < n; ++t) {
    for (int k = 0; k < 4; ++k) s->viol[k] += ws[t].w.viol[k];
    free(ws[t].w.vp); free(ws[t].w.written); free(ws[t].w.logCycle); free(ws[t].w.logMsg);
  }
  pthread_barrier_destroy(&bar);
  free(ws);
  free(th);
  return s->stop;
}
