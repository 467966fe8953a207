// Plan-specialised kernels (lb_rtc.h).
//
// The generated function replays the expression machine of lb_kernels.hip
// (vm_eval, after rust/src/lindblad/eval.rs:791-911) symbolically: every
// stack entry becomes a named register value, and each instruction becomes
// the same helper call (cadd, cmul, dev_powc, ...) on the same operands, so
// the values are bitwise those of the machine.  The library is compiled
// without floating-point contraction, and so is the generated code.
#include "lb_rtc.h"

#include <hip/hiprtc.h>

#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>

extern const int lb_src_count;
extern const char* const lb_src_text[];
extern const char* const lb_src_name[];

namespace {

// A double literal with exactly the bits of v.
std::string lit(double v) {
  char buf[64];
  if (std::isfinite(v)) {
    std::snprintf(buf, sizeof buf, "%a", v);
  } else {
    uint64_t b;
    std::memcpy(&b, &v, 8);
    std::snprintf(buf, sizeof buf, "__hiloint2double((int)0x%08" PRIx32 "u, (int)0x%08" PRIx32 "u)",
                  (uint32_t)(b >> 32), (uint32_t)b);
  }
  return buf;
}

}  // namespace

std::string lb_rtc_coef_source(const LbProgram& prog, const std::vector<int32_t>& expr_dyn, int K,
                               std::string& why) {
  std::ostringstream o;
  o << "__device__ __forceinline__ void lb_gen_coef(dcpx* sl, dcpx* cf, double t) {\n";
  int nv = 0;
  const int nc = prog.n_compounds;
  for (int e = 0; e < nc + K; ++e) {
    if (!expr_dyn[e]) continue;
    std::vector<std::string> st;  // the stack, bottom first
    auto pop = [&]() {
      std::string s = st.back();
      st.pop_back();
      return s;
    };
    o << "  {\n";
    for (int p = prog.expr_start[e]; p < prog.expr_start[e + 1]; ++p) {
      const DevInstr& q = prog.instr[p];
      const std::string v = "v" + std::to_string(nv++);
      o << "    ";
      if (q.op >= 1 && q.op <= 4) {
        o << "const dcpx " << v << " = ";
        if (q.op == 1) o << "mk(" << lit(q.re) << ", " << lit(q.im) << ")";
        else if (q.op == 2) o << "sl[" << q.index << "]";
        else o << "mk(t, 0.0)";
        o << ";\n";
        st.push_back(v);
        continue;
      }
      const size_t need = q.op == 12 || q.op == 13 ? (size_t)std::max(q.argc, 1)
                        : (q.op == 10 || q.op == 11 || q.op == LB_OP_POWC) ? 1 : 2;
      if (st.size() < need) {
        why = "stack underflow in expression " + std::to_string(e);
        return "";
      }
      switch (q.op) {
        case 5: case 6: case 7: case 8: case 9: {
          static const char* fn[] = {"cadd", "csub", "cmul", "cdiv", "dev_cpow"};
          const std::string b = pop(), a = pop();
          o << "const dcpx " << v << " = " << fn[q.op - 5] << "(" << a << ", " << b << ");\n";
          break;
        }
        case 10: {
          const std::string a = pop();
          o << "const dcpx " << v << " = mk(-" << a << ".re, -" << a << ".im);\n";
          break;
        }
        case 11: {
          const std::string a = pop();
          o << "const dcpx " << v << " = mk(" << a << ".re, -" << a << ".im);\n";
          break;
        }
        case 12: {  // builtin on the first argument; exp of a real value inline
          const std::string a = st[st.size() - need];
          st.resize(st.size() - need);
          if (q.function == 4)
            o << "const dcpx " << v << " = " << a << ".im == 0.0 ? mk(exp(" << a << ".re), 0.0) : dev_builtin(4, " << a
              << ");\n";
          else
            o << "const dcpx " << v << " = dev_builtin(" << q.function << ", " << a << ");\n";
          break;
        }
        case 13: {
          // a long tuple argument list stays on the expression machine (its
          // array would live in the kernel's private memory)
          if (need > 16) {
            why = "helper " + std::to_string(q.function) + " with " + std::to_string(need) +
                  " arguments has no generated form";
            return "";
          }
          const std::string arr = "a" + v;
          o << "const dcpx " << arr << "[" << need << "] = {";
          for (size_t j = st.size() - need; j < st.size(); ++j) o << (j + need > st.size() ? ", " : "") << st[j];
          o << "};\n    const dcpx " << v << " = dev_helper(" << q.function << ", " << arr << ", " << need << ");\n";
          st.resize(st.size() - need);
          break;
        }
        case LB_OP_POWC: {
          const std::string a = pop();
          o << "const dcpx " << v << " = dev_powc(" << a << ", " << lit(q.re) << ", " << lit(q.im) << ", "
            << lit(q.aux) << ");\n";
          break;
        }
        case 14: case 15: case 16: case 17: case 18: case 19: {
          static const char* cmp[] = {">", ">=", "<", "<=", "==", "!="};
          const std::string b = pop(), a = pop();
          o << "const dcpx " << v << " = mk((" << a << ".re " << cmp[q.op - 14] << " " << b
            << ".re) ? 1.0 : 0.0, 0.0);\n";
          break;
        }
        default:
          why = "instruction " + std::to_string(q.op) + " has no generated form";
          return "";
      }
      st.push_back(v);
    }
    const std::string r = st.empty() ? "mk(0.0, 0.0)" : st.back();
    if (e < nc) o << "    sl[" << prog.compound_slot[e] << "] = " << r << ";\n";
    else o << "    cf[" << (e - nc) << "] = " << r << ";\n";
    o << "  }\n";
  }
  o << "  cf[" << K << "] = mk(1.0, 0.0);\n}\n";
  return o.str();
}

namespace {

// Code objects compiled in this process, by generated source and kernel
// name: plans with the same programs (several plans of one system) compile once.
struct CodeObject {
  std::vector<char> code;
  std::string lowered;
};
std::mutex g_code_mu;
std::map<std::string, CodeObject> g_code;

int compile_code(const std::string& coef_source, const std::string& name_expr, CodeObject* out,
                 std::string& why) {
  // No variant tables (no part matches), the generated function after the
  // kernels that call it.
  const std::string src = "#define LB_PART 99\n#define LB_NPART 100\n#define LB_GEN_COEF 1\n"
                          "#include \"lb_kernels.hip\"\n" + coef_source;
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "lb_plan_kernel.hip", lb_src_count, lb_src_text, lb_src_name) !=
      HIPRTC_SUCCESS) {
    why = "hiprtcCreateProgram failed";
    return -1;
  }
  hiprtcAddNameExpression(prog, name_expr.c_str());
  std::vector<std::string> words = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  if (const char* extra = std::getenv("CENTREX_LB_RTC_OPTIONS")) {  // diagnostic hook
    std::istringstream in(extra);
    for (std::string w; in >> w;) words.push_back(w);
  }
  std::vector<const char*> opts;
  for (const auto& w : words) opts.push_back(w.c_str());
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
  if (rc != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    why = std::string("runtime compilation failed: ") + hiprtcGetErrorString(rc) + "\n" + log.substr(0, 4000);
    hiprtcDestroyProgram(&prog);
    return -1;
  }
  const char* lowered = nullptr;
  size_t size = 0;
  if (hiprtcGetLoweredName(prog, name_expr.c_str(), &lowered) != HIPRTC_SUCCESS || !lowered ||
      hiprtcGetCodeSize(prog, &size) != HIPRTC_SUCCESS || size == 0) {
    why = "runtime compilation produced no kernel " + name_expr;
    hiprtcDestroyProgram(&prog);
    return -1;
  }
  out->lowered = lowered;
  out->code.resize(size);
  hiprtcGetCode(prog, out->code.data());
  hiprtcDestroyProgram(&prog);
  return 0;
}

}  // namespace

int lb_rtc_compile(const std::string& coef_source, const std::string& name_expr, hipModule_t* module,
                   hipFunction_t* function, std::string& why) {
  const char* extra = std::getenv("CENTREX_LB_RTC_OPTIONS");
  const std::string key = std::string(extra ? extra : "") + "\n" + name_expr + "\n" + coef_source;
  CodeObject obj;
  {
    std::lock_guard<std::mutex> lock(g_code_mu);
    auto it = g_code.find(key);
    if (it != g_code.end()) obj = it->second;
  }
  if (obj.code.empty()) {
    if (compile_code(coef_source, name_expr, &obj, why)) return -1;
    std::lock_guard<std::mutex> lock(g_code_mu);
    g_code.emplace(key, obj);
  }
  const std::string& name = obj.lowered;
  hipError_t e = hipModuleLoadData(module, obj.code.data());
  if (e != hipSuccess) {
    why = std::string("hipModuleLoadData: ") + hipGetErrorString(e);
    return -1;
  }
  e = hipModuleGetFunction(function, *module, name.c_str());
  if (e != hipSuccess) {
    why = std::string("hipModuleGetFunction: ") + hipGetErrorString(e);
    (void)hipModuleUnload(*module);
    *module = nullptr;
    return -1;
  }
  return 0;
}
