// extern "C" boundary of the resident evaluation server (include/sgpp_hip.h).

#include <cstring>
#include <memory>
#include <string>

#include "../../include/sgpp_hip.h"
#include "server.hpp"

struct sgpp_hip_server {
  std::unique_ptr<sgpp_hip::Server> s;
};

namespace sgpp_hip {
// the thread-local message behind sgpp_hip_last_error() (capi.cpp)
void set_last_error(const std::string& msg);
}

namespace {

struct ErrorSink {
  ErrorSink& operator=(const std::string& m) {
    sgpp_hip::set_last_error(m);
    return *this;
  }
};
ErrorSink g_server_error;

template <typename F>
sgpp_hip_status guarded(F&& f) {
  try {
    sgpp_hip::set_last_error("");
    return f();
  } catch (const sgpp_hip::input_error& e) {
    g_server_error = e.what();
    return SGPP_HIP_ERR_INVALID_ARGUMENT;
  } catch (const sgpp_hip::device_error& e) {
    g_server_error = e.what();
    return SGPP_HIP_ERR_DEVICE;
  } catch (const std::exception& e) {
    g_server_error = e.what();
    return SGPP_HIP_ERR_INTERNAL;
  } catch (...) {
    g_server_error = "unknown exception";
    return SGPP_HIP_ERR_INTERNAL;
  }
}
}  // namespace

extern "C" {

sgpp_hip_status sgpp_hip_server_create(const sgpp_hip_server_config* c, sgpp_hip_server** out) {
  return guarded([&] {
    if (!out) throw sgpp_hip::input_error("out is null");
    sgpp_hip::ServerConfig cfg;
    if (c) {
      cfg.device = c->device;
      if (c->precision != SGPP_HIP_PREC_FP64 && c->precision != SGPP_HIP_PREC_FP32_FP64ACC)
        throw sgpp_hip::input_error("unknown precision");
      cfg.precision = c->precision == SGPP_HIP_PREC_FP64 ? sgpp_hip::Precision::FP64
                                                          : sgpp_hip::Precision::FP32;
      if (c->memory_budget) cfg.memory_budget = c->memory_budget;
      if (c->max_batch_queries) cfg.max_batch_queries = c->max_batch_queries;
      if (c->max_items) cfg.max_items = c->max_items;
      if (c->max_partial_doubles) cfg.max_partial_doubles = c->max_partial_doubles;
      if (c->max_inflight > 0) cfg.max_inflight = c->max_inflight;
      if (c->max_requests_per_batch > 0) cfg.max_requests_per_batch = c->max_requests_per_batch;
      cfg.blocks = c->blocks;
      cfg.use_graph = c->use_graph != 0;
      if (c->item_work > 0) cfg.item_work = c->item_work;
      cfg.autotune = c->autotune != 0;
      if (c->autotune_cache) cfg.autotune_cache = c->autotune_cache;
      cfg.algorithm = c->algorithm;
    }
    auto* h = new sgpp_hip_server();
    try {
      h->s.reset(new sgpp_hip::Server(cfg));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_server_destroy(sgpp_hip_server* h) {
  return guarded([&] {
    delete h;
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_server_admit_model(sgpp_hip_server* h, int basis, size_t dim, size_t N,
                                            const uint32_t* level, const uint32_t* index,
                                            const double* alpha, int* model_id) {
  return guarded([&] {
    if (!h || !level || !index || !alpha || !model_id)
      throw sgpp_hip::input_error("null argument");
    if (basis != SGPP_HIP_BASIS_LINEAR && basis != SGPP_HIP_BASIS_MODLINEAR)
      throw sgpp_hip::input_error("unknown basis");
    try {
      *model_id = h->s->admit_model(
          basis == SGPP_HIP_BASIS_LINEAR ? sgpp_hip::Basis::Linear : sgpp_hip::Basis::ModLinear,
          dim, N, level, index, alpha);
    } catch (const sgpp_hip::input_error& e) {
      g_server_error = e.what();
      if (std::strstr(e.what(), "memory budget")) return SGPP_HIP_ERR_BUDGET;
      return SGPP_HIP_ERR_INVALID_ARGUMENT;
    }
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_server_evict_model(sgpp_hip_server* h, int model_id) {
  return guarded([&] {
    if (!h) throw sgpp_hip::input_error("null argument");
    h->s->evict_model(model_id);
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_server_memory_used(sgpp_hip_server* h, size_t* bytes) {
  return guarded([&] {
    if (!h || !bytes) throw sgpp_hip::input_error("null argument");
    *bytes = h->s->memory_used();
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_server_submit(sgpp_hip_server* h, int model_id, size_t q,
                                       const double* queries, double deadline_s,
                                       uint64_t* request_id) {
  return guarded([&] {
    if (!h || !queries || !request_id) throw sgpp_hip::input_error("null argument");
    *request_id = h->s->submit(model_id, q, queries, deadline_s);
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_server_poll(sgpp_hip_server* h, uint64_t request_id, double* out,
                                     size_t out_capacity, double wait_s, int* status,
                                     double* latency_s) {
  return guarded([&] {
    if (!h || !status) throw sgpp_hip::input_error("null argument");
    const int st = h->s->poll(request_id, out, out_capacity, wait_s, latency_s);
    if (st < 0) {
      g_server_error = "unknown request id";
      return SGPP_HIP_ERR_NOT_FOUND;
    }
    *status = st;
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_server_cancel(sgpp_hip_server* h, uint64_t request_id) {
  return guarded([&] {
    if (!h) throw sgpp_hip::input_error("null argument");
    if (!h->s->cancel(request_id)) {
      g_server_error = "request is not queued";
      return SGPP_HIP_ERR_NOT_FOUND;
    }
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_server_inject_error(sgpp_hip_server* h) {
  return guarded([&] {
    if (!h) throw sgpp_hip::input_error("null argument");
    h->s->inject_kernel_error();
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_server_drain(sgpp_hip_server* h) {
  return guarded([&] {
    if (!h) throw sgpp_hip::input_error("null argument");
    h->s->drain();
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_server_stats(sgpp_hip_server* h, sgpp_hip_server_statistics* out) {
  return guarded([&] {
    if (!h || !out) throw sgpp_hip::input_error("null argument");
    const sgpp_hip::ServerStats st = h->s->stats();
    out->requests_done = st.requests_done;
    out->requests_failed = st.requests_failed;
    out->queries_done = st.queries_done;
    out->batches = st.batches;
    out->kernel_launches = st.kernel_launches;
    out->busy_seconds = st.busy_seconds;
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_server_tuned_shape(sgpp_hip_server* h, int basis, int dim, int* found,
                                            sgpp_hip_server_shape* out) {
  return guarded([&] {
    if (!h || !found) throw sgpp_hip::input_error("null argument");
    if (basis != SGPP_HIP_BASIS_LINEAR && basis != SGPP_HIP_BASIS_MODLINEAR)
      throw sgpp_hip::input_error("unknown basis");
    sgpp_hip::TunedShape t;
    *found = h->s->tuned_shape(
                 basis == SGPP_HIP_BASIS_LINEAR ? sgpp_hip::Basis::Linear : sgpp_hip::Basis::ModLinear,
                 dim, &t)
                 ? 1
                 : 0;
    if (*found && out) {
      out->chunk = t.chunk;
      out->cost = t.cost;
    }
    return SGPP_HIP_OK;
  });
}

sgpp_hip_status sgpp_hip_server_model_algorithm(sgpp_hip_server* h, int model_id,
                                                int* algorithm) {
  return guarded([&] {
    if (!h || !algorithm) throw sgpp_hip::input_error("null argument");
    *algorithm = h->s->model_algorithm(model_id) == sgpp_hip::kServeSubspace
                     ? SGPP_HIP_SERVE_SUBSPACE
                     : SGPP_HIP_SERVE_STREAMING;
    return SGPP_HIP_OK;
  });
}

}  // extern "C"
