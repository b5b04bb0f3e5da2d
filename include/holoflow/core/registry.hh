// Factory registry keyed by node kind string (src/holoflow/include/holoflow/core/
// registry.hh:24-68).  register_hip_nodes() installs the kinds of the Doppler-moment path
// under the names the reference's Manager uses (src/holovibes/src/pipeline/
// manager.cc:159-232), plus the fused "DopplerMoments" kind.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "holoflow/core/tasks.hh"

namespace holoflow::core {

class Registry {
 public:
  void register_sync(const std::string &kind, std::unique_ptr<ISyncTaskFactory> f) {
    if (sync_.count(kind) || async_.count(kind)) throw std::invalid_argument("kind already registered: " + kind);
    sync_[kind] = std::move(f);
  }
  void register_async(const std::string &kind, std::unique_ptr<IAsyncTaskFactory> f) {
    if (sync_.count(kind) || async_.count(kind)) throw std::invalid_argument("kind already registered: " + kind);
    async_[kind] = std::move(f);
  }
  const ISyncTaskFactory *sync_factory(const std::string &kind) const {
    auto it = sync_.find(kind);
    return it == sync_.end() ? nullptr : it->second.get();
  }
  const IAsyncTaskFactory *async_factory(const std::string &kind) const {
    auto it = async_.find(kind);
    return it == async_.end() ? nullptr : it->second.get();
  }
  const ITaskFactory &factory(const std::string &kind) const {
    if (auto *s = sync_factory(kind)) return *s;
    if (auto *a = async_factory(kind)) return *a;
    throw std::invalid_argument("unknown node kind: " + kind);
  }
  std::vector<std::string> kinds() const {
    std::vector<std::string> k;
    for (const auto &kv : sync_) k.push_back(kv.first);
    for (const auto &kv : async_) k.push_back(kv.first);
    return k;
  }

 private:
  std::map<std::string, std::unique_ptr<ISyncTaskFactory>> sync_;
  std::map<std::string, std::unique_ptr<IAsyncTaskFactory>> async_;
};

// Registers every node kind this build implements.
void register_hip_nodes(Registry &registry);

}  // namespace holoflow::core
