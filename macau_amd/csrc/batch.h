// batch.h -- many independent models advanced as one launch set (batch.cpp).
#pragma once
#include <vector>

#include "model.h"

namespace mgpu {

class ModelBatch {
 public:
  explicit ModelBatch(std::vector<Model*> models);  // not owned; each initialised
  ~ModelBatch();
  ModelBatch(const ModelBatch&) = delete;
  ModelBatch& operator=(const ModelBatch&) = delete;
  void iterate(int n);  // n Gibbs iterations of every model, asynchronous
  void synchronize();   // wait for all; raise sampler failures
  std::vector<Model*> models;
  bool use_graph = true;

 private:
  void launch_all();
  hipStream_t root = nullptr;
  std::vector<hipEvent_t> ev_fork, ev_join;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool launched_any = false, last_graph = true;
};

}  // namespace mgpu
