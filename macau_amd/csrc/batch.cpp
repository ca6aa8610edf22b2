// batch.cpp -- many independent models advanced as one launch set.
//
// The reference runs one model per process (MacauX::run, macau.cpp:94-148).
// A batch holds N initialised models (any sizes and latent dimensions) and
// advances all of them by one Gibbs iteration per replay of ONE HIP graph:
// the graph is captured once from a root stream that forks into every
// model's own streams (each model's iteration is its step_launch: sampling,
// prior updates on its side stream, noise, evaluation) and joins them again,
// so the instances' kernels are independent branches the GPU can run side by
// side.  Each model keeps its own device state, draws and iteration counter,
// so a model advanced in a batch gives the same chain as the same model
// advanced alone.  Without the batch graph (use_graph off) each model
// replays its own iteration graph on its own streams, one after another on
// the host and side by side on the device.
#include "batch.h"

namespace mgpu {

ModelBatch::ModelBatch(std::vector<Model*> ms) : models(std::move(ms)) {
  if (models.empty()) throw std::runtime_error("a batch needs at least one model");
  for (Model* m : models) {
    if (!m || !m->initialized) throw std::runtime_error("every model of a batch must be initialised (macau_init)");
    for (auto& p : m->priors) {
      const auto* mp = dynamic_cast<const MacauPrior*>(p.get());
      if (mp && !mp->capturable())
        throw std::runtime_error("a batch cannot hold a Macau prior solved by block CG (its host reads convergence "
                                 "flags inside the iteration)");
    }
  }
  HIPCHK(hipStreamCreateWithFlags(&root, hipStreamNonBlocking));
  ev_fork.resize(models.size());
  ev_join.resize(models.size());
  for (size_t q = 0; q < models.size(); q++) {
    HIPCHK(hipEventCreateWithFlags(&ev_fork[q], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_join[q], hipEventDisableTiming));
  }
}

ModelBatch::~ModelBatch() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (auto e : ev_fork) (void)hipEventDestroy(e);
  for (auto e : ev_join) (void)hipEventDestroy(e);
  if (root) (void)hipStreamDestroy(root);
}

void ModelBatch::launch_all() {
  // every model waits for the root, runs its iteration on its own streams,
  // and the root waits for every model
  for (size_t q = 0; q < models.size(); q++) {
    HIPCHK(hipEventRecord(ev_fork[q], root));
    HIPCHK(hipStreamWaitEvent(models[q]->stream, ev_fork[q], 0));
    models[q]->step_launch();
    HIPCHK(hipEventRecord(ev_join[q], models[q]->stream));
    HIPCHK(hipStreamWaitEvent(root, ev_join[q], 0));
  }
}

void ModelBatch::iterate(int n) {
  if (n > 0 && launched_any && use_graph != last_graph) {
    // switching between the two launch forms: the new form's streams wait
    // for the work the other form queued
    for (size_t q = 0; q < models.size(); q++) {
      if (use_graph) {
        HIPCHK(hipEventRecord(ev_join[q], models[q]->stream));
        HIPCHK(hipStreamWaitEvent(root, ev_join[q], 0));
      } else {
        HIPCHK(hipEventRecord(ev_fork[q], root));
        HIPCHK(hipStreamWaitEvent(models[q]->stream, ev_fork[q], 0));
      }
    }
  }
  if (n > 0) {
    launched_any = true;
    last_graph = use_graph;
  }
  if (!use_graph) {
    // every model replays its own iteration (its own graph) on its own
    // streams; nothing orders the models, so the hardware queues run them
    // side by side
    for (int i = 0; i < n; i++)
      for (Model* m : models) m->iterate(1);
    return;
  }
  for (int i = 0; i < n; i++) {
    if (!exec) {
      HIPCHK(hipStreamBeginCapture(root, hipStreamCaptureModeThreadLocal));
      launch_all();
      HIPCHK(hipStreamEndCapture(root, &graph));
      HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(exec, root));
    for (Model* m : models) m->host_iter++;
  }
}

void ModelBatch::synchronize() {
  HIPCHK(hipStreamSynchronize(root));
  for (Model* m : models) {
    m->sync();
    m->check_fail();
  }
}

}  // namespace mgpu
