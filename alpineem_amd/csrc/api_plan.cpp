// api_plan.cpp -- host-only part of the libaem C ABI (include/aem.h):
// error reporting and the setup stage (aem_plan_*).
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "../../include/aem.h"
#include "api_internal.hpp"
#include "plan.hpp"

namespace {
thread_local std::string g_err;
}

namespace aem {
void set_error(const std::string& m) { g_err = m; }
}  // namespace aem

extern "C" {

const char* aem_last_error(void) { return g_err.c_str(); }

const char* aem_version(void) { return "libaem 0.1 (gfx950)"; }

void aem_plan_options_init(aem_plan_options* opt) {
    if (opt) aem::default_options(*opt);
}

aem_status aem_plan_from_text_ex(const char* deck_text, const char* base_dir, const aem_plan_options* opt,
                                 aem_plan** out) {
    if (!deck_text || !out) {
        aem::set_error("aem_plan_from_text: null argument");
        return AEM_ERR_STATE;
    }
    *out = nullptr;
    return aem::guard([&]() {
        aem_plan_options o;
        aem::default_options(o);
        if (opt) o = *opt;
        aem::Plan* p = aem::build_plan(deck_text, base_dir ? base_dir : ".", o);
        *out = reinterpret_cast<aem_plan*>(p);
    });
}

aem_status aem_plan_from_text(const char* deck_text, const char* base_dir, aem_plan** out) {
    return aem_plan_from_text_ex(deck_text, base_dir, nullptr, out);
}

aem_status aem_plan_from_inputs(const char* inputs_path, aem_plan** out) {
    return aem_plan_from_inputs_ex(inputs_path, nullptr, out);
}

aem_status aem_plan_from_inputs_ex(const char* inputs_path, const aem_plan_options* opt, aem_plan** out) {
    if (!inputs_path || !out) {
        aem::set_error("aem_plan_from_inputs: null argument");
        return AEM_ERR_STATE;
    }
    *out = nullptr;
    std::ifstream f(inputs_path, std::ios::binary);
    if (!f) {
        aem::set_error(std::string("cannot open deck '") + inputs_path + "'");
        return AEM_ERR_IO;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    std::string path(inputs_path);
    size_t slash = path.find_last_of('/');
    std::string dir = slash == std::string::npos ? "." : path.substr(0, slash);
    return aem_plan_from_text_ex(ss.str().c_str(), dir.c_str(), opt, out);
}

const aem_plan_view* aem_plan_view_get(const aem_plan* plan) {
    if (!plan) return nullptr;
    return &reinterpret_cast<const aem::Plan*>(plan)->v;
}

aem_status aem_plan_write_geometry(const aem_plan* plan, const char* path) {
    if (!plan || !path) {
        aem::set_error("aem_plan_write_geometry: null argument");
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() { aem::write_geometry(*reinterpret_cast<const aem::Plan*>(plan), path); });
}

void aem_plan_destroy(aem_plan* plan) { delete reinterpret_cast<aem::Plan*>(plan); }

}  // extern "C"
