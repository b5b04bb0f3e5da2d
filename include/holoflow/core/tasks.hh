// Task and factory interfaces of the holoflow dataflow runtime, HIP-typed.
// Mirrors src/holoflow/include/holoflow/core/tasks.hh: SyncCtx (:81-90), OpResult
// (:111-116), ITask owned-slot hooks acquire_input/release_output (:136-151), ISyncTask
// (:182-190), IAsyncTask try_push/try_pop (:208-220), InPlace/TaskKind/InferResult
// (:222-250), SyncCreateCtx{stream} (:252-255), factories infer/create/update (:260-334).
// The only type change is the stream: hipStream_t instead of cudaStream_t.  The event
// mailboxes of the UI control plane (holoflow_event) are not part of this build; the
// context keeps the cancellation flag.
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <memory>
#include <optional>
#include <span>
#include <vector>

#include "holoflow/core/json.hh"
#include "holoflow/core/tensor.hh"

namespace holoflow::core {

struct SyncCtx {
  std::span<TView> inputs;
  std::span<TView> outputs;
  std::atomic<bool> *cancelled = nullptr;
};

struct AsyncPushCtx {
  std::span<TView> inputs;
  std::atomic<bool> *cancelled = nullptr;
};

struct AsyncPopCtx {
  std::span<TView> outputs;
  std::atomic<bool> *cancelled = nullptr;
};

enum class OpResult : uint8_t { Ok, NotReady, Cancelled, Eof };

class IOStorageAccess {
 public:
  virtual ~IOStorageAccess() = default;
  virtual Storage &owned_input_storage(size_t index) = 0;
  virtual Storage &owned_output_storage(size_t index) = 0;
};

class ITask {
 public:
  virtual ~ITask() = default;
  // Owned-slot hooks of asynchronous tasks (ring queues); sync tasks keep the defaults.
  virtual std::optional<TView> acquire_input(int) { return std::nullopt; }
  virtual void release_output(int) {}
  void bind_storage_access(IOStorageAccess *sa) { storage_access_ = sa; }

 protected:
  IOStorageAccess &storage_access() { return *storage_access_; }

 private:
  IOStorageAccess *storage_access_ = nullptr;
};

class ISyncTask : public ITask {
 public:
  virtual OpResult execute(SyncCtx &ctx) = 0;
};

class IAsyncTask : public ITask {
 public:
  virtual OpResult try_push(AsyncPushCtx &ctx) = 0;
  virtual OpResult try_pop(AsyncPopCtx &ctx) = 0;
};

struct InPlace {
  int in_idx;
  int out_idx;
};

enum class TaskKind { Sync, Async };

struct InferResult {
  std::vector<TDesc> input_descs;
  std::vector<TDesc> output_descs;
  std::vector<InPlace> in_place;
  std::vector<bool> owned_inputs;
  std::vector<bool> owned_outputs;
  TaskKind kind = TaskKind::Sync;
};

struct SyncCreateCtx {
  hipStream_t stream = nullptr;
};

struct AsyncCreateCtx {
  hipStream_t producer_stream = nullptr;
  hipStream_t consumer_stream = nullptr;
};

class ITaskFactory {
 public:
  virtual ~ITaskFactory() = default;
  virtual InferResult infer(std::span<const TDesc> input_descs, const Json &settings) const = 0;
};

class ISyncTaskFactory : public ITaskFactory {
 public:
  virtual std::unique_ptr<ISyncTask> create(std::span<const TDesc> input_descs, const Json &settings,
                                            const SyncCreateCtx &ctx) const = 0;
  // Default update recreates the task (tasks.hh:292-305); factories override it to reuse.
  virtual std::unique_ptr<ISyncTask> update(std::unique_ptr<ISyncTask> old_task,
                                            std::span<const TDesc> input_descs, const Json &settings,
                                            const SyncCreateCtx &ctx) const {
    (void)old_task;
    return create(input_descs, settings, ctx);
  }
};

class IAsyncTaskFactory : public ITaskFactory {
 public:
  virtual std::unique_ptr<IAsyncTask> create(std::span<const TDesc> input_descs, const Json &settings,
                                             const AsyncCreateCtx &ctx) const = 0;
  virtual std::unique_ptr<IAsyncTask> update(std::unique_ptr<IAsyncTask> old_task,
                                             std::span<const TDesc> input_descs, const Json &settings,
                                             const AsyncCreateCtx &ctx) const {
    (void)old_task;
    return create(input_descs, settings, ctx);
  }
};

}  // namespace holoflow::core
