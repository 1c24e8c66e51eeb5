#include "runtime/timer.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <iomanip>
#include <memory>
#include <mutex>
#include <sstream>
#include <vector>

namespace spla_amd {
namespace timer {
namespace {

using Clock = std::chrono::steady_clock;

struct Node {
  std::string name;
  std::vector<double> timings;
  std::vector<std::unique_ptr<Node>> children;  // in order of first appearance

  Node* child(const std::string& n) {
    for (auto& c : children)
      if (c->name == n) return c.get();
    children.push_back(std::make_unique<Node>());
    children.back()->name = n;
    return children.back().get();
  }
  double total() const {
    double s = 0.0;
    for (double t : timings) s += t;
    return s;
  }
};

struct Registry {
  std::mutex mutex;
  Node root;
  std::vector<std::string> warnings;
  uint64_t generation = 1;  // bumped by clear(): open scopes of older generations are dropped
};

Registry& registry() {
  static Registry r;
  return r;
}

struct Open {
  Node* node;
  Clock::time_point begin;
};

struct ThreadStack {
  uint64_t generation = 0;
  std::vector<Open> open;
};

thread_local ThreadStack tStack;

std::atomic<bool> gLibraryScopes{false};

// Keeps the thread's open scopes only if they belong to the current tree.
void sync_generation(const Registry& r) {
  if (tStack.generation != r.generation) {
    tStack.open.clear();
    tStack.generation = r.generation;
  }
}

std::string escaped(const std::string& s) {
  std::string out;
  for (char c : s) {
    if (c == '"' || c == '\\') {
      out += '\\';
      out += c;
    } else if (static_cast<unsigned char>(c) < 0x20) {
      char buf[8];
      std::snprintf(buf, sizeof(buf), "\\u%04x", c);
      out += buf;
    } else {
      out += c;
    }
  }
  return out;
}

void write_json(const std::vector<std::unique_ptr<Node>>& nodes, const std::string& pad,
                std::ostringstream& os) {
  os << "{\n";
  const std::string inner = pad + "  ";
  const std::string field = inner + "  ";
  for (size_t i = 0; i < nodes.size(); ++i) {
    const Node& n = *nodes[i];
    os << inner << '"' << escaped(n.name) << "\" : {\n" << field << "\"timings\" : [";
    for (size_t j = 0; j < n.timings.size(); ++j) os << (j ? ", " : "") << n.timings[j];
    os << "],\n" << field << "\"sub-timings\" : ";
    write_json(n.children, field, os);
    os << inner << '}' << (i + 1 < nodes.size() ? "," : "") << '\n';
  }
  os << pad << "}\n";
}

double median(std::vector<double> v) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const size_t h = v.size() / 2;
  return v.size() % 2 ? v[h] : 0.5 * (v[h - 1] + v[h]);
}

struct Row {
  std::string label;
  const Node* node;
  double topTotal;
  double parentTotal;
};

void collect(const Node& n, int depth, double topTotal, double parentTotal,
             std::vector<Row>& rows) {
  const double total = n.total();
  const double top = depth == 0 ? total : topTotal;
  rows.push_back({std::string(2 * depth, ' ') + n.name, &n, top, parentTotal});
  for (const auto& c : n.children) collect(*c, depth + 1, top, total, rows);
}

}  // namespace

void start(const std::string& name) {
  Registry& r = registry();
  std::lock_guard<std::mutex> lock(r.mutex);
  sync_generation(r);
  Node* parent = tStack.open.empty() ? &r.root : tStack.open.back().node;
  Node* node = parent->child(name);
  tStack.open.push_back({node, Clock::now()});
}

void stop(const std::string& name) {
  const Clock::time_point end = Clock::now();
  Registry& r = registry();
  std::lock_guard<std::mutex> lock(r.mutex);
  sync_generation(r);
  if (tStack.open.empty() || tStack.open.back().node->name != name) {
    r.warnings.push_back("timer: stop(\"" + name +
                         "\") does not match the innermost open scope; ignored");
    return;
  }
  const Open o = tStack.open.back();
  tStack.open.pop_back();
  o.node->timings.push_back(std::chrono::duration<double>(end - o.begin).count());
}

std::string json() {
  Registry& r = registry();
  std::lock_guard<std::mutex> lock(r.mutex);
  std::ostringstream os;
  os << std::scientific;
  write_json(r.root.children, "", os);
  return os.str();
}

std::string table() {
  Registry& r = registry();
  std::lock_guard<std::mutex> lock(r.mutex);
  std::vector<Row> rows;
  for (const auto& c : r.root.children) collect(*c, 0, 0.0, 0.0, rows);
  size_t labelWidth = 8;
  for (const Row& row : rows) labelWidth = std::max(labelWidth, row.label.size() + 3);
  const char* headers[] = {"#", "Total", "%", "Parent %", "Median", "Min", "Max"};
  const int widths[] = {8, 14, 10, 10, 14, 14, 14};
  size_t totalWidth = labelWidth;
  for (int w : widths) totalWidth += w;

  std::ostringstream os;
  for (const std::string& w : r.warnings) os << w << '\n';
  os << std::string(totalWidth, '=') << '\n' << std::setw(static_cast<int>(labelWidth)) << "";
  for (int i = 0; i < 7; ++i) os << std::right << std::setw(widths[i]) << headers[i];
  os << '\n' << std::string(totalWidth, '-') << '\n';
  for (const Row& row : rows) {
    const Node& n = *row.node;
    const double total = n.total();
    const double mn = n.timings.empty() ? 0.0 : *std::min_element(n.timings.begin(), n.timings.end());
    const double mx = n.timings.empty() ? 0.0 : *std::max_element(n.timings.begin(), n.timings.end());
    const double pct = row.topTotal > 0.0 ? 100.0 * total / row.topTotal : 100.0;
    const double ppct = row.parentTotal > 0.0 ? 100.0 * total / row.parentTotal : 100.0;
    os << std::left << std::setw(static_cast<int>(labelWidth)) << row.label << std::right
       << std::setw(widths[0]) << n.timings.size() << std::scientific << std::setprecision(4)
       << std::setw(widths[1]) << total << std::fixed << std::setprecision(2)
       << std::setw(widths[2]) << pct << std::setw(widths[3]) << ppct << std::scientific
       << std::setprecision(4) << std::setw(widths[4]) << median(n.timings)
       << std::setw(widths[5]) << mn << std::setw(widths[6]) << mx << '\n';
  }
  os << std::string(totalWidth, '=') << '\n';
  return os.str();
}

void clear() {
  Registry& r = registry();
  std::lock_guard<std::mutex> lock(r.mutex);
  r.root.children.clear();
  r.warnings.clear();
  ++r.generation;
}

void set_library_scopes(bool enable) { gLibraryScopes.store(enable); }
bool library_scopes() { return gLibraryScopes.load(std::memory_order_relaxed); }

}  // namespace timer
}  // namespace spla_amd
