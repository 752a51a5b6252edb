// mlp_capi.cpp -- C ABI of the MLP surrogate network.

#include <hip/hip_runtime.h>

#include <memory>

#include "handles.hpp"
#include "mlp.hpp"

using namespace nosh;

struct nosh_mlp_s {
  Mlp net;
  DBuf<float> x, t, y;
};

namespace {
template <class F>
int mguard(F&& f) {
  try {
    f();
    set_last_error("");
    return NOSH_OK;
  } catch (const nosh::Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return NOSH_ERR_INVALID;
  }
}
}  // namespace

extern "C" {

int nosh_mlp_create(int n_in, int width, int depth, int n_out, uint64_t seed, nosh_mlp_t* out) {
  return mguard([&] {
    NOSH_CHECK(out, kInvalid, "null argument");
    auto h = std::make_unique<nosh_mlp_s>();
    h->net.init(n_in, width, depth, n_out, seed);
    *out = h.release();
  });
}

int nosh_mlp_num_params(nosh_mlp_t mlp, int64_t* nparams, int64_t* nweights) {
  return mguard([&] {
    NOSH_CHECK(mlp, kInvalid, "null mlp");
    if (nparams) *nparams = mlp->net.nparams;
    if (nweights) *nweights = mlp->net.nweights;
  });
}

int nosh_mlp_get_params(nosh_mlp_t mlp, float* params) {
  return mguard([&] {
    NOSH_CHECK(mlp && params, kInvalid, "null argument");
    mlp->net.master.download(params, mlp->net.nparams, default_stream());
    NOSH_HIP(hipStreamSynchronize(default_stream()));
  });
}

int nosh_mlp_set_params(nosh_mlp_t mlp, const float* params) {
  return mguard([&] {
    NOSH_CHECK(mlp && params, kInvalid, "null argument");
    mlp->net.master.upload(params, mlp->net.nparams, default_stream());
    mlp->net.sync_bf16(default_stream());
    NOSH_HIP(hipStreamSynchronize(default_stream()));
  });
}

int nosh_mlp_forward_device(nosh_mlp_t mlp, const float* X, int64_t M, float* Y, void* stream) {
  return mguard([&] {
    NOSH_CHECK(mlp && X && Y && M >= 1, kInvalid, "invalid argument");
    mlp->net.forward(X, M, Y, stream ? static_cast<hipStream_t>(stream) : default_stream());
  });
}

int nosh_mlp_forward(nosh_mlp_t mlp, const float* X, int64_t M, float* Y) {
  return mguard([&] {
    NOSH_CHECK(mlp && X && Y && M >= 1, kInvalid, "invalid argument");
    Mlp& n = mlp->net;
    hipStream_t s = default_stream();
    if (mlp->x.n < size_t(M * n.nin)) mlp->x.alloc(M * n.nin);
    if (mlp->y.n < size_t(M * n.nout)) mlp->y.alloc(M * n.nout);
    mlp->x.upload(X, M * n.nin, s);
    n.forward(mlp->x.p, M, mlp->y.p, s);
    mlp->y.download(Y, M * n.nout, s);
    NOSH_HIP(hipStreamSynchronize(s));
  });
}

int nosh_mlp_infer(nosh_mlp_t mlp, const float* X, int64_t M, float* Y) {
  return mguard([&] {
    NOSH_CHECK(mlp && X && Y && M >= 1, kInvalid, "invalid argument");
    Mlp& n = mlp->net;
    hipStream_t s = default_stream();
    if (mlp->x.n < size_t(M * n.nin)) mlp->x.alloc(M * n.nin);
    if (mlp->y.n < size_t(M * n.nout)) mlp->y.alloc(M * n.nout);
    mlp->x.upload(X, M * n.nin, s);
    n.infer(mlp->x.p, M, mlp->y.p, s);
    mlp->y.download(Y, M * n.nout, s);
    NOSH_HIP(hipStreamSynchronize(s));
  });
}

int nosh_mlp_infer_device(nosh_mlp_t mlp, const float* X, int64_t M, float* Y, void* stream) {
  return mguard([&] {
    NOSH_CHECK(mlp && X && Y && M >= 1, kInvalid, "invalid argument");
    mlp->net.infer(X, M, Y, stream ? static_cast<hipStream_t>(stream) : default_stream());
  });
}

int nosh_mlp_gradients(nosh_mlp_t mlp, const float* X, const float* T, int64_t M, float* grad,
                       double* loss) {
  return mguard([&] {
    NOSH_CHECK(mlp && X && T && M >= 1, kInvalid, "invalid argument");
    Mlp& n = mlp->net;
    hipStream_t s = default_stream();
    if (mlp->x.n < size_t(M * n.nin)) mlp->x.alloc(M * n.nin);
    if (mlp->t.n < size_t(M * n.nout)) mlp->t.alloc(M * n.nout);
    mlp->x.upload(X, M * n.nin, s);
    mlp->t.upload(T, M * n.nout, s);
    const double l = n.gradients(mlp->x.p, mlp->t.p, M, s);
    if (loss) *loss = l;
    if (grad) {
      n.grad.download(grad, n.nparams, s);
      NOSH_HIP(hipStreamSynchronize(s));
    }
  });
}

int nosh_mlp_train_step_device(nosh_mlp_t mlp, const float* X, const float* T, int64_t M,
                               float lr, double* loss, void* stream) {
  return mguard([&] {
    NOSH_CHECK(mlp && X && T && M >= 1, kInvalid, "invalid argument");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : default_stream();
    const double l = mlp->net.gradients(X, T, M, s);
    mlp->net.adam(lr, s);
    if (loss) *loss = l;
  });
}

int nosh_mlp_destroy(nosh_mlp_t mlp) {
  return mguard([&] { delete mlp; });
}

}  // extern "C"
