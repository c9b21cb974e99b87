/*
 * Multi-threaded restatement of the reference plugin's CPU platform -- TEST INFRASTRUCTURE ONLY.
 *
 * Timed by bench.py's cpu_baseline leg on the host cores (the reference CPU platform itself needs OpenMM, which is
 * not available).  It follows the algorithm of the CPU platform, not its code:
 *
 *   pair forces (CpuInteraxPairForceFvec.h:170-394): each pair force separately, in FP32, with its own neighbour
 *   list built on every call (atoms sorted into cells and cut into blocks of 8; for each block the atoms within the
 *   cutoff of its bounding box, each pair once), blocks taken by the threads from a shared counter, the 8 lanes of a
 *   block evaluated together against one neighbour at a time (FP32 minimum image with rounding, blended branches:
 *   every lane computes, the mask selects), forces summed per thread and combined at the end;
 *
 *   three-body (CpuInteraxThreeBodyForce.cpp:89-188): in FP64, site centroids, then for every site A the sites
 *   B > A of other molecules within the cutoff (all pairs, as the CPU platform does), then the triples A < B < C
 *   found by marking the neighbours of A, site forces per thread, shared equally among a site's particles.
 *
 * The FP64 geometry helpers and the ATM triple of the oracle are reused (this file includes it).  Built into
 * oracle/libcpuplatform.so with OpenMP; the HIP product never loads it.
 */
#include "interax_oracle.c"

#include <omp.h>
#include <stdint.h>

#define CPL_BLOCK 8

/* ------------------------------------------------------------------------------------------------------------ */
/* FP32 pair terms (InteraxPairTerms.h:167-337 in single precision)                                              */
/* ------------------------------------------------------------------------------------------------------------ */

typedef struct {
    int order;
    int about_smaller_width;
    float max_delta, max_t, max_tx;
} cpl_scheme;

static float TAYLOR_F[ORC_TAYLOR_ORDER+1][ORC_TAYLOR_TERMS], TAYLOR_DF[ORC_TAYLOR_ORDER+1][ORC_TAYLOR_TERMS];

static void init_tables(void) {
    for (int n = 0; n <= ORC_TAYLOR_ORDER; n++)
        for (int k = 0; k < ORC_TAYLOR_TERMS; k++) {
            TAYLOR_F[n][k] = (float) OVERLAP_V[n][k];
            TAYLOR_DF[n][k] = (float) overlap_d(n, k);
        }
}

static void scheme_f(int method, cpl_scheme* s) {
    orc_taylor_scheme d;
    if (method == 2)
        orc_scheme_order2(&d);
    else
        orc_scheme_accurate(0, &d);
    s->order = d.order;
    s->about_smaller_width = d.about_smaller_width;
    s->max_delta = (float) d.max_delta;
    s->max_t = (float) d.max_t;
    s->max_tx = (float) fmin(d.max_tx, 1e30);
}

static inline void overlap_f(float r, float a, float b, const cpl_scheme* s, float* S, float* dS) {
    const float l = (s->about_smaller_width ? fminf(a, b) : a);
    const float invL = 1.0f/l;
    const float delta = a+b-l-l;
    const float t = delta*invL;
    const float x = r*invL;
    const float absT = fabsf(t);
    /* both branches, then the blend (the CPU platform evaluates every lane and selects) */
    float sumValue = 0, sumDeriv = 0;
    for (int n = s->order; n >= 0; n--) {
        float value = 0, dv = 0;
        for (int k = ORC_TAYLOR_TERMS-1; k >= 0; k--) {
            value = value*x + TAYLOR_F[n][k];
            dv = dv*x + TAYLOR_DF[n][k];
        }
        sumValue = sumValue*t + value;
        sumDeriv = sumDeriv*t + dv;
    }
    const float prefactor = expf(-x)*invL*invL*invL*(float) (1.0/M_PI);
    const float taylorS = prefactor*sumValue, taylorD = prefactor*invL*sumDeriv;
    const float aa = a*a, bb = b*b;
    const float d = bb-aa;
    const float invD2 = 1.0f/(d*d);
    const float expA = expf(-r/a), expB = expf(-r/b);
    const float c = 4*aa*bb*invD2/d;
    const float f = c*(expA-expB) + r*invD2*(a*expA+b*expB);
    const float dfdr = c*(expB/b-expA/a) + invD2*(a*expA+b*expB-r*(expA+expB));
    const float invR = 1.0f/r;
    const float closedS = f*invR*(float) (1.0/(8*M_PI));
    const float closedD = (r*dfdr-f)*invR*invR*(float) (1.0/(8*M_PI));
    const int taylor = (fabsf(delta) < s->max_delta) || ((absT < s->max_t) && (absT*x < s->max_tx));
    *S = (taylor ? taylorS : closedS);
    *dS = (taylor ? taylorD : closedD);
}

/* Tang-Toennies f_6, f_8, f_10 (series tail to k = 40, switch at x = (n+1)/2) */
static inline void tang_toennies_f(float x, float damping[3], float dd[3]) {
    const float expX = expf(-x);
    float term = 1, partial = 1;
    float terms[11], partials[11];
    terms[0] = term;
    partials[0] = partial;
    for (int k = 1; k <= 10; k++) {
        term = term*x*(1.0f/k);
        partial = partial+term;
        terms[k] = term;
        partials[k] = partial;
    }
    float tail = 0;
    for (int k = 11; k <= 40; k++) {
        term = term*x*(1.0f/k);
        tail = tail+term;
    }
    for (int i = 0; i < 3; i++) {
        const int n = 6+2*i;
        float series = tail;
        for (int k = n+1; k <= 10; k++)
            series = series+terms[k];
        damping[i] = (x < 0.5f*(n+1) ? expX*series : 1-expX*partials[n]);
        dd[i] = expX*terms[n];
    }
}

/* Energy and dE/dr of one pair force; p1, p2 rows of FP32 parameters (first = lower index for overlap terms). */
static inline void pair_term_f(int kind, float r, const float* p1, const float* p2, float beta, const cpl_scheme* s, float* e,
                               float* dEdR) {
    if (kind == ORC_DISPERSION) {
        const float prefactor = p1[3]*p2[3]/fmaxf(p1[3]*p2[0]+p2[3]*p1[0], 1e-30f);
        float c[3];
        c[0] = prefactor*p1[0]*p2[0];
        c[1] = 1.5f*prefactor*(p1[0]*p2[1]+p1[1]*p2[0]);
        c[2] = prefactor*(2*(p1[0]*p2[2]+p1[2]*p2[0])+4.2f*p1[1]*p2[1]);
        const float dxdr = 2*beta/(p1[4]+p2[4]);
        float f[3], df[3];
        tang_toennies_f(dxdr*r, f, df);
        const float invR = 1.0f/r;
        const float invR2 = invR*invR;
        float invRn = invR2*invR2*invR2;
        float en = 0, d = 0;
        for (int i = 0; i < 3; i++) {
            const float n = (float) (6+2*i);
            en -= f[i]*c[i]*invRn;
            d -= c[i]*invRn*(df[i]*dxdr - n*f[i]*invR);
            invRn *= invR2;
        }
        *e = en;
        *dEdR = d;
    }
    else if (kind == ORC_EXCHANGE) {
        float S, dS;
        overlap_f(r, p1[1], p2[1], s, &S, &dS);
        *e = p1[0]*p2[0]*S;
        *dEdR = p1[0]*p2[0]*dS;
    }
    else {
        float S, dS;
        overlap_f(r, p1[2], p2[2], s, &S, &dS);
        const float widthSum = p1[2]+p2[2];
        const float norm = (float) (8*M_PI)*widthSum*widthSum*widthSum;
        const float St = norm*S;
        const float k = 0.5f*(p1[0]+p2[0]), l = 0.5f*(p1[1]+p2[1]);
        *e = (k+l*St)*St;
        *dEdR = (k+2*l*St)*norm*dS;
    }
}

static inline int gate_f(int kind, const float* p) {
    return (kind == ORC_DISPERSION ? p[3] != 0 : kind == ORC_EXCHANGE ? p[1] != 0 : p[2] != 0);
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Pair force: neighbour blocks and the block loop                                                               */
/* ------------------------------------------------------------------------------------------------------------ */

typedef struct {
    int n, kind, periodic, cutoff_method, rs;
    float cutoff2, beta, lambda, softcoreShift, switchDistance, invSwitchWidth;
    int useSwitch, useSoftcore;
    float box[3], invBox[3];
    const float* pos;        /* 4 per atom, original order */
    const float* params;     /* rs per atom */
    const unsigned char* alchemical;
    const int* sortedAtom;   /* atoms in cell order, padded to a multiple of 8 with -1 */
    const int* sortedIndex;  /* position of each atom in sortedAtom */
    const exclusion_table* excl;
    cpl_scheme scheme;
    /* cell grid of the sorted atoms */
    int ncell[3];
    float cellSize[3], lo[3];
    const int* cellStart;    /* ranges in sortedAtom per cell */
    const int* cellEnd;
} pair_setup;

static inline int cell_of(float p, int k, const pair_setup* s) {
    float u = (p-s->lo[k])/s->cellSize[k];
    int c = (int) floorf(u);
    if (s->periodic) {
        c %= s->ncell[k];
        if (c < 0)
            c += s->ncell[k];
    }
    else {
        if (c < 0)
            c = 0;
        if (c >= s->ncell[k])
            c = s->ncell[k]-1;
    }
    return c;
}

/* One block of 8 sorted atoms against the atoms within the cutoff of its bounding box that come later in the sorted
   order (pairs inside the block: the later lane). */
static void block_ixn(const pair_setup* s, int block, float* force, double* energy, long long* pairs, int* scratch, int scratchCap) {
    int atom[CPL_BLOCK];
    float x[CPL_BLOCK], y[CPL_BLOCK], z[CPL_BLOCK];
    float bp[CPL_BLOCK][5];
    float alch[CPL_BLOCK];
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    int valid = 0;
    for (int l = 0; l < CPL_BLOCK; l++) {
        atom[l] = s->sortedAtom[CPL_BLOCK*block+l];
        const int a = (atom[l] < 0 ? 0 : atom[l]);
        x[l] = s->pos[4*a];
        y[l] = s->pos[4*a+1];
        z[l] = s->pos[4*a+2];
        for (int k = 0; k < s->rs; k++)
            bp[l][k] = s->params[s->rs*a+k];
        alch[l] = (s->alchemical != NULL && s->alchemical[a] ? 1.0f : 0.0f);
        if (atom[l] >= 0) {
            valid++;
            lo[0] = fminf(lo[0], x[l]); hi[0] = fmaxf(hi[0], x[l]);
            lo[1] = fminf(lo[1], y[l]); hi[1] = fmaxf(hi[1], y[l]);
            lo[2] = fminf(lo[2], z[l]); hi[2] = fmaxf(hi[2], z[l]);
        }
    }
    if (valid == 0)
        return;
    /* neighbour atoms: later in the sorted order, within the cutoff of the bounding box */
    int count = 0;
    const float rc = sqrtf(s->cutoff2);
    if (s->cutoff_method == ORC_NO_CUTOFF) {
        for (int t = CPL_BLOCK*block; t < s->n; t++)
            if (count < scratchCap)
                scratch[count++] = t;
    }
    else {
        int c0[3], c1[3];
        for (int k = 0; k < 3; k++) {
            const float p0 = (k == 0 ? lo[0] : k == 1 ? lo[1] : lo[2]), p1 = (k == 0 ? hi[0] : k == 1 ? hi[1] : hi[2]);
            c0[k] = (int) floorf((p0-rc-s->lo[k])/s->cellSize[k]);
            c1[k] = (int) floorf((p1+rc-s->lo[k])/s->cellSize[k]);
            if (!s->periodic) {
                if (c0[k] < 0) c0[k] = 0;
                if (c1[k] >= s->ncell[k]) c1[k] = s->ncell[k]-1;
            }
            else if (c1[k]-c0[k]+1 > s->ncell[k]) {
                c0[k] = 0;
                c1[k] = s->ncell[k]-1;
            }
        }
        for (int cz = c0[2]; cz <= c1[2]; cz++)
            for (int cy = c0[1]; cy <= c1[1]; cy++)
                for (int cx = c0[0]; cx <= c1[0]; cx++) {
                    int ix = cx, iy = cy, iz = cz;
                    if (s->periodic) {
                        ix = ((ix%s->ncell[0])+s->ncell[0])%s->ncell[0];
                        iy = ((iy%s->ncell[1])+s->ncell[1])%s->ncell[1];
                        iz = ((iz%s->ncell[2])+s->ncell[2])%s->ncell[2];
                    }
                    const int cell = (iz*s->ncell[1]+iy)*s->ncell[0]+ix;
                    for (int t = s->cellStart[cell]; t < s->cellEnd[cell]; t++) {
                        if (t < CPL_BLOCK*block)
                            continue;
                        const int j = s->sortedAtom[t];
                        /* distance from the bounding box (nearest image of the point relative to the box centre) */
                        float d2 = 0;
                        const float pj[3] = {s->pos[4*j], s->pos[4*j+1], s->pos[4*j+2]};
                        for (int k = 0; k < 3; k++) {
                            const float centre = 0.5f*(lo[k]+hi[k]), half = 0.5f*(hi[k]-lo[k]);
                            float dk = pj[k]-centre;
                            if (s->periodic)
                                dk -= s->box[k]*rintf(dk*s->invBox[k]);
                            dk = fmaxf(fabsf(dk)-half, 0.0f);
                            d2 += dk*dk;
                        }
                        if (d2 < s->cutoff2 && count < scratchCap)
                            scratch[count++] = t;
                    }
                }
    }
    float fx[CPL_BLOCK] = {0}, fy[CPL_BLOCK] = {0}, fz[CPL_BLOCK] = {0};
    double blockEnergy = 0;
    for (int q = 0; q < count; q++) {
        const int t = scratch[q];
        const int j = s->sortedAtom[t];
        const float xj = s->pos[4*j], yj = s->pos[4*j+1], zj = s->pos[4*j+2];
        const float* pjp = s->params+s->rs*j;
        const float alchJ = (s->alchemical != NULL && s->alchemical[j] ? 1.0f : 0.0f);
        float lane_fx[CPL_BLOCK], lane_fy[CPL_BLOCK], lane_fz[CPL_BLOCK], lane_e[CPL_BLOCK];
        int include[CPL_BLOCK];
        for (int l = 0; l < CPL_BLOCK; l++) {
            float dx = x[l]-xj, dy = y[l]-yj, dz = z[l]-zj;
            if (s->periodic) {
                dx -= rintf(dx*s->invBox[0])*s->box[0];
                dy -= rintf(dy*s->invBox[1])*s->box[1];
                dz -= rintf(dz*s->invBox[2])*s->box[2];
            }
            const float r2 = dx*dx+dy*dy+dz*dz;
            int inc = (atom[l] >= 0) && (CPL_BLOCK*block+l < t);
            if (inc && s->cutoff_method != ORC_NO_CUTOFF)
                inc = (r2 < s->cutoff2);
            if (inc)
                inc = gate_f(s->kind, bp[l]) && gate_f(s->kind, pjp) && !is_excluded(s->excl, atom[l], j);
            include[l] = inc;
            lane_fx[l] = lane_fy[l] = lane_fz[l] = lane_e[l] = 0;
            if (!inc)
                continue;
            const float invR = 1.0f/sqrtf(r2);
            const float r = r2*invR;
            const int alchemicalLane = fabsf(alch[l]-alchJ) > 0.5f;
            float rEff = r, softcoreFactor = 1;
            if (s->useSoftcore && alchemicalLane) {
                rEff = powf(powf(r, 6.0f)+s->softcoreShift, 1.0f/6.0f);
                softcoreFactor = powf(r/rEff, 5.0f);
            }
            float e, dEdR;
            if (s->kind != ORC_DISPERSION && atom[l] > j)
                pair_term_f(s->kind, rEff, pjp, bp[l], s->beta, &s->scheme, &e, &dEdR);
            else
                pair_term_f(s->kind, rEff, bp[l], pjp, s->beta, &s->scheme, &e, &dEdR);
            dEdR *= softcoreFactor;
            if (s->useSwitch && r > s->switchDistance) {
                const float u = (r-s->switchDistance)*s->invSwitchWidth;
                const float sv = 1+u*u*u*(-10+u*(15-u*6));
                const float sd = u*u*(-30+u*(60-u*30))*s->invSwitchWidth;
                dEdR = dEdR*sv+e*sd;
                e *= sv;
            }
            const float scale = (alchemicalLane ? s->lambda : 1.0f);
            lane_e[l] = scale*e;
            const float forceScale = scale*dEdR*invR;
            lane_fx[l] = dx*forceScale;
            lane_fy[l] = dy*forceScale;
            lane_fz[l] = dz*forceScale;
        }
        float sx = 0, sy = 0, sz = 0;
        for (int l = 0; l < CPL_BLOCK; l++) {
            if (!include[l])
                continue;
            (*pairs)++;
            blockEnergy += lane_e[l];
            fx[l] -= lane_fx[l];
            fy[l] -= lane_fy[l];
            fz[l] -= lane_fz[l];
            sx += lane_fx[l];
            sy += lane_fy[l];
            sz += lane_fz[l];
        }
        force[4*j] += sx;
        force[4*j+1] += sy;
        force[4*j+2] += sz;
    }
    for (int l = 0; l < CPL_BLOCK; l++)
        if (atom[l] >= 0) {
            force[4*atom[l]] += fx[l];
            force[4*atom[l]+1] += fy[l];
            force[4*atom[l]+2] += fz[l];
        }
    *energy += blockEnergy;
}

/*
 * One pair force in FP32 on `threads` threads (all host cores when <= 0).  Forces are ADDED to forces[3*n];
 * energy receives the force's energy (no long range correction); pairs the number of interacting pairs.
 */
int cpl_pair_compute(const orc_pair_force* f, const double* positions, const double* box, double lambda, int threads, double* forces,
                     double* energy, long long* pairs) {
    static int initialized = 0;
    if (!initialized) {
        init_tables();
        initialized = 1;
    }
    if (f == NULL || positions == NULL || forces == NULL || energy == NULL)
        return -1;
    const int n = f->num_particles;
    const int periodic = (f->method == ORC_CUTOFF_PERIODIC);
    if (periodic && !is_rectangular(box))
        return -1;
    if (threads <= 0)
        threads = omp_get_max_threads();
    pair_setup s;
    memset(&s, 0, sizeof(s));
    s.n = n;
    s.kind = f->kind;
    s.rs = row_size(f->kind);
    s.periodic = periodic;
    s.cutoff_method = f->method;
    s.cutoff2 = (float) (f->cutoff*f->cutoff);
    s.beta = (float) f->damping_scale;
    s.lambda = (float) lambda;
    s.useSoftcore = (f->softcore_radius > 0 && any_alchemical(f->alchemical, n));
    s.softcoreShift = (float) ((1-lambda)*pow(f->softcore_radius, 6));
    s.useSwitch = (f->method != ORC_NO_CUTOFF && f->use_switch);
    s.switchDistance = (float) f->switching_distance;
    s.invSwitchWidth = (float) (1/(f->cutoff-f->switching_distance));
    s.alchemical = f->alchemical;
    scheme_f(f->overlap_method, &s.scheme);
    /* FP32 copies of positions (wrapped into the box) and parameters */
    float* pos = (float*) malloc(sizeof(float)*4*(size_t) (n > 0 ? n : 1));
    float* params = (float*) malloc(sizeof(float)*(size_t) s.rs*(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            double p = positions[3*i+k];
            if (periodic)
                p -= floor(p/box[4*k])*box[4*k];
            pos[4*i+k] = (float) p;
        }
        pos[4*i+3] = 0;
        for (int k = 0; k < s.rs; k++)
            params[s.rs*i+k] = (float) f->params[s.rs*i+k];
    }
    s.pos = pos;
    s.params = params;
    /* cell grid of side >= cutoff (one cell without a cutoff) and the atoms in cell order */
    float hiB[3];
    for (int k = 0; k < 3; k++) {
        if (periodic) {
            s.box[k] = (float) box[4*k];
            s.invBox[k] = (float) (1/box[4*k]);
            s.lo[k] = 0;
            hiB[k] = s.box[k];
        }
        else {
            s.lo[k] = 1e30f;
            hiB[k] = -1e30f;
            for (int i = 0; i < n; i++) {
                s.lo[k] = fminf(s.lo[k], pos[4*i+k]);
                hiB[k] = fmaxf(hiB[k], pos[4*i+k]);
            }
            if (n == 0)
                s.lo[k] = hiB[k] = 0;
        }
        const float extent = fmaxf(hiB[k]-s.lo[k], 1e-6f);
        int cells = (f->method == ORC_NO_CUTOFF ? 1 : (int) floorf(extent/(float) f->cutoff));
        if (cells < 1)
            cells = 1;
        if (cells > 256)
            cells = 256;
        s.ncell[k] = cells;
        s.cellSize[k] = (periodic ? s.box[k]/cells : fmaxf(extent/cells, (float) f->cutoff)*(1+1e-6f));
    }
    const int numCells = s.ncell[0]*s.ncell[1]*s.ncell[2];
    int* cellOfAtom = (int*) malloc(sizeof(int)*(size_t) (n > 0 ? n : 1));
    int* cellStart = (int*) calloc((size_t) numCells+1, sizeof(int));
    int* cellEnd = (int*) calloc((size_t) numCells, sizeof(int));
    for (int i = 0; i < n; i++) {
        const int c = (cell_of(pos[4*i+2], 2, &s)*s.ncell[1]+cell_of(pos[4*i+1], 1, &s))*s.ncell[0]+cell_of(pos[4*i], 0, &s);
        cellOfAtom[i] = c;
        cellStart[c+1]++;
    }
    for (int c = 0; c < numCells; c++)
        cellStart[c+1] += cellStart[c];
    const int numBlocks = (n+CPL_BLOCK-1)/CPL_BLOCK;
    int* sortedAtom = (int*) malloc(sizeof(int)*(size_t) (numBlocks > 0 ? numBlocks : 1)*CPL_BLOCK);
    int* fill = (int*) calloc((size_t) numCells, sizeof(int));
    for (int i = 0; i < n; i++) {
        const int c = cellOfAtom[i];
        sortedAtom[cellStart[c]+fill[c]++] = i;
    }
    for (int t = n; t < numBlocks*CPL_BLOCK; t++)
        sortedAtom[t] = -1;
    for (int c = 0; c < numCells; c++)
        cellEnd[c] = cellStart[c+1];
    s.sortedAtom = sortedAtom;
    s.cellStart = cellStart;
    s.cellEnd = cellEnd;
    exclusion_table excl;
    build_exclusions(f, &excl);
    s.excl = &excl;
    /* the block loop: threads take blocks from a shared counter; forces and energy per thread */
    float* threadForce = (float*) calloc((size_t) threads*4*(n > 0 ? n : 1), sizeof(float));
    double* threadEnergy = (double*) calloc((size_t) threads, sizeof(double));
    long long* threadPairs = (long long*) calloc((size_t) threads, sizeof(long long));
    const int scratchCap = n+CPL_BLOCK;
    #pragma omp parallel num_threads(threads)
    {
        const int tid = omp_get_thread_num();
        int* scratch = (int*) malloc(sizeof(int)*(size_t) scratchCap);
        float* force = threadForce+(size_t) tid*4*(n > 0 ? n : 1);
        double e = 0;
        long long p = 0;
        #pragma omp for schedule(dynamic, 1)
        for (int b = 0; b < numBlocks; b++)
            block_ixn(&s, b, force, &e, &p, scratch, scratchCap);
        threadEnergy[tid] = e;
        threadPairs[tid] = p;
        free(scratch);
    }
    double total = 0;
    long long totalPairs = 0;
    for (int t = 0; t < threads; t++) {
        total += threadEnergy[t];
        totalPairs += threadPairs[t];
        const float* force = threadForce+(size_t) t*4*(n > 0 ? n : 1);
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++)
                forces[3*i+k] += force[4*i+k];
    }
    *energy = total;
    if (pairs != NULL)
        *pairs = totalPairs;
    free(threadForce);
    free(threadEnergy);
    free(threadPairs);
    free(excl.offsets);
    free(excl.partners);
    free(fill);
    free(sortedAtom);
    free(cellOfAtom);
    free(cellStart);
    free(cellEnd);
    free(pos);
    free(params);
    return 0;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Three-body (CpuInteraxThreeBodyForce.cpp:89-188)                                                              */
/* ------------------------------------------------------------------------------------------------------------ */

/*
 * Forces ADDED to forces[3*num_particles]; energy receives the energy; triples the number of interacting triples.
 */
int cpl_threebody_compute(const orc_threebody_force* f, int num_particles, const double* positions, const double* box, double lambda,
                          int threads, double* forces, double* energy, long long* triples) {
    if (f == NULL || positions == NULL || forces == NULL || energy == NULL)
        return -1;
    const int n = f->num_sites;
    if (threads <= 0)
        threads = omp_get_max_threads();
    double* sites = (double*) malloc(sizeof(double)*3*(size_t) (n > 0 ? n : 1));
    int* nbrCount = (int*) calloc((size_t) n+1, sizeof(int));
    int** nbr = (int**) calloc((size_t) (n > 0 ? n : 1), sizeof(int*));
    const double cutoff2 = f->cutoff*f->cutoff;
    /* site positions, then the neighbours B > A of other molecules within the cutoff (all pairs) */
    #pragma omp parallel for num_threads(threads) schedule(static)
    for (int i = 0; i < n; i++) {
        double c[3] = {0, 0, 0};
        const int b0 = f->site_offsets[i], b1 = f->site_offsets[i+1];
        for (int q = b0; q < b1; q++)
            for (int k = 0; k < 3; k++)
                c[k] += positions[3*f->site_particles[q]+k];
        for (int k = 0; k < 3; k++)
            sites[3*i+k] = c[k]/(b1-b0);
    }
    #pragma omp parallel num_threads(threads)
    {
        int cap = 64;
        int* list = (int*) malloc(sizeof(int)*(size_t) cap);
        #pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < n; i++) {
            int count = 0;
            const int mi = (f->molecule[i] < 0 ? -1-i : f->molecule[i]);
            for (int j = i+1; j < n; j++) {
                if (mi == (f->molecule[j] < 0 ? -1-j : f->molecule[j]))
                    continue;
                double d[3];
                delta_vec(sites+3*i, sites+3*j, box, f->periodic, d);
                if (dot3(d, d) < cutoff2) {
                    if (count == cap) {
                        cap *= 2;
                        list = (int*) realloc(list, sizeof(int)*(size_t) cap);
                    }
                    list[count++] = j;
                }
            }
            nbr[i] = (int*) malloc(sizeof(int)*(size_t) (count > 0 ? count : 1));
            memcpy(nbr[i], list, sizeof(int)*(size_t) count);
            nbrCount[i] = count;
        }
        free(list);
    }
    /* the triples A < B < C: C a neighbour of both A and B, found by marking the neighbours of A */
    double* threadForce = (double*) calloc((size_t) threads*3*(n > 0 ? n : 1), sizeof(double));
    double* threadEnergy = (double*) calloc((size_t) threads, sizeof(double));
    long long* threadTriples = (long long*) calloc((size_t) threads, sizeof(long long));
    #pragma omp parallel num_threads(threads)
    {
        const int tid = omp_get_thread_num();
        int* marks = (int*) malloc(sizeof(int)*(size_t) (n > 0 ? n : 1));
        for (int i = 0; i < n; i++)
            marks[i] = -1;
        double* siteForce = threadForce+(size_t) tid*3*(n > 0 ? n : 1);
        double e = 0;
        long long count = 0;
        #pragma omp for schedule(dynamic, 1)
        for (int a = 0; a < n; a++) {
            for (int q = 0; q < nbrCount[a]; q++)
                marks[nbr[a][q]] = a;
            for (int q = 0; q < nbrCount[a]; q++) {
                const int b = nbr[a][q];
                for (int r = 0; r < nbrCount[b]; r++) {
                    const int c = nbr[b][r];
                    if (marks[c] != a)
                        continue;
                    double c9 = f->scale_factor*orc_c9(f->alpha[a], f->frequency[a], f->alpha[b], f->frequency[b], f->alpha[c],
                                                      f->frequency[c]);
                    if (f->alchemical != NULL && (f->alchemical[a] || f->alchemical[b] || f->alchemical[c]))
                        c9 *= lambda;
                    if (c9 == 0)
                        continue;
                    double ab[3], ac[3], bc[3];
                    delta_vec(sites+3*a, sites+3*b, box, f->periodic, ab);
                    delta_vec(sites+3*a, sites+3*c, box, f->periodic, ac);
                    delta_vec(sites+3*b, sites+3*c, box, f->periodic, bc);
                    double rates[3] = {2*f->damping_scale/(f->width[a]+f->width[b]), 2*f->damping_scale/(f->width[a]+f->width[c]),
                                       2*f->damping_scale/(f->width[b]+f->width[c])};
                    e += orc_triple(ab, ac, bc, c9, f->damping_method == 1 ? rates : NULL, siteForce+3*a, siteForce+3*b, siteForce+3*c,
                                    f->centroid_damping_radius);
                    count++;
                }
            }
        }
        threadEnergy[tid] = e;
        threadTriples[tid] = count;
        free(marks);
    }
    double total = 0;
    long long totalTriples = 0;
    for (int t = 0; t < threads; t++) {
        total += threadEnergy[t];
        totalTriples += threadTriples[t];
    }
    for (int i = 0; i < n; i++) {
        double fsum[3] = {0, 0, 0};
        for (int t = 0; t < threads; t++)
            for (int k = 0; k < 3; k++)
                fsum[k] += threadForce[(size_t) t*3*(n > 0 ? n : 1)+3*i+k];
        const int b0 = f->site_offsets[i], b1 = f->site_offsets[i+1];
        for (int q = b0; q < b1; q++)
            for (int k = 0; k < 3; k++)
                forces[3*f->site_particles[q]+k] += fsum[k]/(b1-b0);
    }
    *energy = total;
    if (triples != NULL)
        *triples = totalTriples;
    for (int i = 0; i < n; i++)
        free(nbr[i]);
    free(nbr);
    free(nbrCount);
    free(threadForce);
    free(threadEnergy);
    free(threadTriples);
    free(sites);
    (void) num_particles;
    return 0;
}

int cpl_num_threads(void) {
    return omp_get_max_threads();
}
