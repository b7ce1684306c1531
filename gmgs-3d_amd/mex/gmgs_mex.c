/*
 * MATLAB MEX gateway for libgmgs (include/gmgs.h).
 *
 * Build where MATLAB is installed:
 *   mex -R2018a gmgs_mex.c -I../../include -L.. -lgmgs
 *
 * Calls (h is a uint64 handle):
 *   h = gmgs_mex('create', voxelizedVolume_, boundingBox_, params)
 *         params fields (all optional): moduls_, poissonRatio_, tol_, maxIT_,
 *         weightFactorJacobi_, coarsestResolutionControl_, numLevels, precision, flags
 *   gmgs_mex('set_bc', h, fixingCond_, loadingCond_)
 *   [U_, iterationHist_, info] = gmgs_mex('solve', h)
 *   S = gmgs_mex('stress', h, U_)          % cartesianStressField_ (numNodes x 6)
 *   y = gmgs_mex('atx', h, x, level)       % @AtX handle
 *   z = gmgs_mex('vcycle', h, r)           % @Vcycle handle
 *   gmgs_mex('destroy', h)
 * Errors map to mexErrMsgIdAndTxt('gmgs:<status>', ...); a non-converged
 * solve maps to the reference's warning (CG_solver.m:52-55).
 */
#include <string.h>

#include "gmgs.h"
#include "mex.h"

static gmgs_ctx* handle_of(const mxArray* a) {
  if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1) mexErrMsgIdAndTxt("gmgs:handle", "bad handle");
  return (gmgs_ctx*)(uintptr_t)(*(uint64_t*)mxGetData(a));
}

static void check(gmgs_ctx* c, int st) {
  char id[64];
  if (st == GMGS_OK) return;
  if (st == GMGS_ERR_NOT_CONVERGED) {
    mexWarnMsgIdAndTxt("gmgs:notConverged", "Exceed the maximum iterate numbers");
    return;
  }
  snprintf(id, sizeof(id), "gmgs:status%d", st);
  mexErrMsgIdAndTxt(id, "%s: %s", gmgs_status_string(st), gmgs_last_error(c));
}

static double field_or(const mxArray* s, const char* name, double dflt) {
  const mxArray* f;
  if (!s || !mxIsStruct(s)) return dflt;
  f = mxGetField(s, 0, name);
  return (f && !mxIsEmpty(f)) ? mxGetScalar(f) : dflt;
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  char cmd[32];
  if (nrhs < 1 || mxGetString(prhs[0], cmd, sizeof(cmd)) != 0)
    mexErrMsgIdAndTxt("gmgs:usage", "first argument must be a command string");

  if (!strcmp(cmd, "create")) {
    gmgs_params p;
    gmgs_ctx* c = NULL;
    const mwSize* d;
    int32_t* vox;
    mwSize n, i;
    const double* src;
    int st;
    if (nrhs < 3) mexErrMsgIdAndTxt("gmgs:usage", "create(vox, bbox, params)");
    gmgs_default_params(&p);
    p.youngs_modulus = field_or(nrhs > 3 ? prhs[3] : NULL, "moduls_", p.youngs_modulus);
    p.poisson_ratio = field_or(nrhs > 3 ? prhs[3] : NULL, "poissonRatio_", p.poisson_ratio);
    p.tol = field_or(nrhs > 3 ? prhs[3] : NULL, "tol_", p.tol);
    p.max_it = (int32_t)field_or(nrhs > 3 ? prhs[3] : NULL, "maxIT_", p.max_it);
    p.jacobi_weight = field_or(nrhs > 3 ? prhs[3] : NULL, "weightFactorJacobi_", p.jacobi_weight);
    p.coarsest_resolution_control =
        field_or(nrhs > 3 ? prhs[3] : NULL, "coarsestResolutionControl_", p.coarsest_resolution_control);
    p.num_levels = (int32_t)field_or(nrhs > 3 ? prhs[3] : NULL, "numLevels", 0);
    p.precision = (int32_t)field_or(nrhs > 3 ? prhs[3] : NULL, "precision", GMGS_PREC_FP64);
    p.flags = (int32_t)field_or(nrhs > 3 ? prhs[3] : NULL, "flags", 0);
    if (mxGetNumberOfDimensions(prhs[1]) != 3) mexErrMsgIdAndTxt("gmgs:usage", "vox must be 3-D");
    d = mxGetDimensions(prhs[1]);
    n = d[0] * d[1] * d[2];
    vox = (int32_t*)mxMalloc(n * sizeof(int32_t));
    if (mxIsInt32(prhs[1])) {
      memcpy(vox, mxGetData(prhs[1]), n * sizeof(int32_t));
    } else {
      src = mxGetPr(prhs[1]);
      for (i = 0; i < n; ++i) vox[i] = (int32_t)src[i];
    }
    {
      /* boundingBox_ is 2x3 column-major: [xmin ymin zmin; xmax ymax zmax] */
      double bb[6];
      const double* b = mxGetPr(prhs[2]);
      bb[0] = b[0]; bb[1] = b[2]; bb[2] = b[4];
      bb[3] = b[1]; bb[4] = b[3]; bb[5] = b[5];
      st = gmgs_create(&p, vox, (int64_t)d[0], (int64_t)d[1], (int64_t)d[2], bb, &c);
    }
    mxFree(vox);
    check(NULL, st);
    plhs[0] = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
    *(uint64_t*)mxGetData(plhs[0]) = (uint64_t)(uintptr_t)c;
    return;
  }
  if (nrhs < 2) mexErrMsgIdAndTxt("gmgs:usage", "missing handle");
  {
    gmgs_ctx* c = handle_of(prhs[1]);
    if (!strcmp(cmd, "destroy")) {
      gmgs_destroy(c);
    } else if (!strcmp(cmd, "set_bc")) {
      mwSize k = mxGetNumberOfElements(prhs[2]), i;
      int32_t* fx = (int32_t*)mxMalloc(k * sizeof(int32_t));
      const double* f = mxGetPr(prhs[2]);
      for (i = 0; i < k; ++i) fx[i] = (int32_t)f[i];
      check(c, gmgs_set_bc(c, fx, (int64_t)k, mxGetPr(prhs[3]), (int64_t)mxGetM(prhs[3])));
      mxFree(fx);
    } else if (!strcmp(cmd, "solve")) {
      int64_t nn = 0, maxit = 0, cap, rows, i;
      gmgs_solve_info info;
      double *hist, *H;
      gmgs_query(c, GMGS_Q_NUM_NODES, 1, &nn, 1);
      check(c, gmgs_query(c, GMGS_Q_MAX_IT, 1, &maxit, 1));
      plhs[0] = mxCreateDoubleMatrix((mwSize)(3 * nn), 1, mxREAL);
      /* iterationHist_ has at most maxIT_ + 1 rows (CG_solver.m:17,49) */
      cap = maxit + 2;
      hist = (double*)mxMalloc(sizeof(double) * 3 * (size_t)cap);
      check(c, gmgs_solve(c, mxGetPr(plhs[0]), &info, hist, cap));
      rows = info.hist_len < cap ? info.hist_len : cap;
      if (nlhs > 1) {
        plhs[1] = mxCreateDoubleMatrix((mwSize)rows, 3, mxREAL);
        H = mxGetPr(plhs[1]);
        for (i = 0; i < rows; ++i) {
          H[i] = hist[3 * i];
          H[i + rows] = hist[3 * i + 1];
          H[i + 2 * rows] = hist[3 * i + 2];
        }
      }
      if (nlhs > 2) {
        const char* names[] = {"its", "converged", "relres", "true_relres", "solve_seconds"};
        plhs[2] = mxCreateStructMatrix(1, 1, 5, names);
        mxSetField(plhs[2], 0, "its", mxCreateDoubleScalar(info.its));
        mxSetField(plhs[2], 0, "converged", mxCreateDoubleScalar(info.converged));
        mxSetField(plhs[2], 0, "relres", mxCreateDoubleScalar(info.relres));
        mxSetField(plhs[2], 0, "true_relres", mxCreateDoubleScalar(info.true_relres));
        mxSetField(plhs[2], 0, "solve_seconds", mxCreateDoubleScalar(info.solve_seconds));
      }
      mxFree(hist);
    } else if (!strcmp(cmd, "stress")) {
      int64_t nn = 0;
      gmgs_query(c, GMGS_Q_NUM_NODES, 1, &nn, 1);
      plhs[0] = mxCreateDoubleMatrix((mwSize)nn, 6, mxREAL);
      check(c, gmgs_stress(c, nrhs > 2 ? mxGetPr(prhs[2]) : NULL, mxGetPr(plhs[0])));
    } else if (!strcmp(cmd, "atx")) {
      int level = nrhs > 3 ? (int)mxGetScalar(prhs[3]) : 1;
      plhs[0] = mxCreateDoubleMatrix(mxGetM(prhs[2]), 1, mxREAL);
      check(c, gmgs_atx(c, level, mxGetPr(prhs[2]), mxGetPr(plhs[0])));
    } else if (!strcmp(cmd, "vcycle")) {
      plhs[0] = mxCreateDoubleMatrix(mxGetM(prhs[2]), 1, mxREAL);
      check(c, gmgs_vcycle(c, mxGetPr(prhs[2]), mxGetPr(plhs[0])));
    } else {
      mexErrMsgIdAndTxt("gmgs:usage", "unknown command %s", cmd);
    }
  }
}
