// The pybind11 binding a maintainer of the reference adds to its _core
// module (python/igaos_bindings.cpp) for the batched engine, which has no
// counterpart there: `add_batch_bindings(m)` registers
//
//   solve_batch(models, precision=32, tolerance=1e-4, time_limit=120.0,
//               max_iterations=500000, mode="lockstep") -> list of dicts
//
// over the reference's own io::Model objects (src/io/model.hpp), through the
// C-ABI of libigaos_pdhg.so (include/igaos_pdhg.h): the instances' arrays
// are concatenated in order, uploaded once by igaos_batch_create and solved
// together by igaos_batch_run.  Each result holds status (the reference's
// status strings), objective, pinf, dinf, rel_gap, iterations, x and y.
// Errors (no GPU, malformed models) raise RuntimeError with the library's
// message.
//
// Built as its own extension module (IGAOS_BATCH_STANDALONE) it also
// exposes read_mps (the reference's parse_mps), so it can be compiled and
// loaded without the rest of the reference's module.
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "igaos_pdhg.h"
#include "model.hpp"

namespace py = pybind11;

namespace {

py::list solve_batch(const std::vector<const igaos::io::Model*>& models, int precision,
                     double tolerance, double time_limit, long max_iterations,
                     const std::string& mode) {
    if (mode != "lockstep" && mode != "persistent")
        throw std::invalid_argument("mode must be 'lockstep' or 'persistent'");
    const int count = static_cast<int>(models.size());
    std::vector<int> m, n, nnz, ap, ai, cp, ci;
    std::vector<double> c, cl, cu, rmin, rmax, ax, acx, oc;
    for (const igaos::io::Model* md : models) {
        if (!md) throw std::invalid_argument("None in the model list");
        m.push_back(md->m);
        n.push_back(md->n);
        nnz.push_back(static_cast<int>(md->ax.size()));
        auto cat = [](auto& dst, const auto& src) { dst.insert(dst.end(), src.begin(), src.end()); };
        cat(c, md->c);
        cat(cl, md->cl);
        cat(cu, md->cu);
        cat(rmin, md->rmin);
        cat(rmax, md->rmax);
        cat(ap, md->ap);  // instance-local offsets, as the reference stores them
        cat(ai, md->ai);
        cat(ax, md->ax);
        cat(cp, md->cp);
        cat(ci, md->ci);
        cat(acx, md->acx);
        oc.push_back(md->obj_const);
    }
    std::unique_ptr<igaos_batch, void (*)(igaos_batch*)> b(
        igaos_batch_create(count, m.data(), n.data(), nnz.data(), c.data(), cl.data(),
                           cu.data(), rmin.data(), rmax.data(), ap.data(), ai.data(),
                           ax.data(), cp.data(), ci.data(), acx.data(), oc.data(), precision),
        igaos_batch_destroy);
    if (!b) throw std::runtime_error(igaos_batch_last_error());
    if (igaos_batch_set_mode(b.get(), mode == "lockstep" ? 1 : 0) != 0)
        throw std::runtime_error(igaos_batch_last_error());
    igaos_options opt;
    igaos_options_default(&opt);
    opt.tolerance = tolerance;
    opt.time_limit_s = time_limit;
    opt.max_iterations = max_iterations;
    std::vector<double> x(c.size()), y(rmin.size());
    std::vector<igaos_batch_result> res(count);
    {
        py::gil_scoped_release nogil;
        if (igaos_batch_run(b.get(), &opt, x.data(), y.data(), res.data()) != 0)
            throw std::runtime_error(igaos_batch_last_error());
    }
    py::list out;
    size_t xo = 0, yo = 0;
    for (int i = 0; i < count; ++i) {
        py::dict d;
        d["status"] = igaos_status_name(res[i].status);
        d["objective"] = res[i].objective;
        d["pinf"] = res[i].pinf;
        d["dinf"] = res[i].dinf;
        d["rel_gap"] = res[i].rel_gap;
        d["iterations"] = res[i].iterations;
        d["x"] = std::vector<double>(x.begin() + xo, x.begin() + xo + n[i]);
        d["y"] = std::vector<double>(y.begin() + yo, y.begin() + yo + m[i]);
        xo += n[i];
        yo += m[i];
        out.append(d);
    }
    return out;
}

}  // namespace

void add_batch_bindings(py::module_& m) {
    m.def("solve_batch", &solve_batch, py::arg("models"), py::arg("precision") = 32,
          py::arg("tolerance") = 1e-4, py::arg("time_limit") = 120.0,
          py::arg("max_iterations") = 500000L, py::arg("mode") = "lockstep");
}

#ifdef IGAOS_BATCH_STANDALONE
PYBIND11_MODULE(igaos_batch_ext, m) {
    py::class_<igaos::io::Model>(m, "Model")
        .def_readonly("m", &igaos::io::Model::m)
        .def_readonly("n", &igaos::io::Model::n);
    m.def("read_mps", [](const std::string& path) {
        auto md = std::make_unique<igaos::io::Model>();
        igaos::io::parse_mps(path, *md);
        return md;
    });
    add_batch_bindings(m);
}
#endif
