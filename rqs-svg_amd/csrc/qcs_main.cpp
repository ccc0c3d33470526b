// bin/qcs: runs a circuit plug-in (a shared object exporting circuit_init / circuit_run)
// on libqcs.so, once per sample, printing one JSON line per sample.
//
// Behaviour and output follow the reference driver (src/qcs_main.cpp:64-201, options of
// src/qcs_args.ggo:8-11): -s/--num-samples N, --mapping CSV (physical -> logical),
// -r/--reversed-mapping (use the inverse of the mapping), --output-statevector FILE; each
// sample records the circuit between two events and prints
//   {"sample_num": i, "clbits": "<clbit n-1 ... clbit 0>", "elapsed_time": seconds}
// and, between samples, resets mapping, state, clbits and measurement lists.
// Additions: --seed N (reproducible measurements, qcs_ext_set_seed), --precision MODE and
// --block-qubits K (qcs_ext_set_precision), and a NULL check on the plug-in symbols.
// The option parser is getopt_long instead of generated gengetopt code.
#include <qcs.h>
#include <qcs_ext.h>

#include <dlfcn.h>
#include <getopt.h>
#include <unistd.h>

#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
    std::vector<bit_num_t> parse_mapping(const std::string &text)
    {
        if (text.empty())
            throw std::runtime_error("mapping string must not be empty");
        std::vector<bit_num_t> out;
        std::stringstream ss(text);
        std::string tok;
        while (std::getline(ss, tok, ','))
        {
            if (tok.empty())
                throw std::runtime_error("mapping contains an empty entry");
            char *end = nullptr;
            errno = 0;
            const long v = std::strtol(tok.c_str(), &end, 10);
            if (errno != 0 || end != tok.c_str() + tok.size() || v < SHRT_MIN || v > SHRT_MAX)
                throw std::runtime_error("mapping entry is not a qubit number: " + tok);
            out.push_back((bit_num_t)v);
        }
        return out;
    }

    // perm_p2l -> perm_l2p with the same validation as set_mapping
    std::vector<bit_num_t> invert(const std::vector<bit_num_t> &p2l, int n)
    {
        if ((int)p2l.size() != n)
            throw std::runtime_error("mapping size " + std::to_string(p2l.size()) + " does not match num_qubits " + std::to_string(n));
        std::vector<bit_num_t> l2p(n, -1);
        for (int p = 0; p < n; ++p)
        {
            int l = p2l[p];
            if (l < 0 || l >= n)
                throw std::runtime_error("mapping value " + std::to_string(l) + " is out of range [0, " + std::to_string(n) + ")");
            if (l2p[l] != -1)
                throw std::runtime_error("mapping value " + std::to_string(l) + " appears multiple times");
            l2p[l] = (bit_num_t)p;
        }
        return l2p;
    }

    std::string last_error;
    void on_error(const char *msg, size_t len) { last_error.assign(msg, len); }

    void must(int rc, const char *what)
    {
        if (rc != 1)
            throw std::runtime_error(std::string(what) + " failed" + (last_error.empty() ? "" : ": " + last_error));
    }

    void usage(FILE *f)
    {
        std::fprintf(f, "usage: qcs [OPTIONS] USER_CIRCUIT_SO\n"
                        "  -s, --num-samples N          number of measurement samples (default 1)\n"
                        "      --mapping CSV            physical-to-logical qubit mapping\n"
                        "  -r, --reversed-mapping       use the inverse of the mapping\n"
                        "      --output-statevector F   write the final state vector to F\n"
                        "      --seed N                 seed of the measurement engine\n"
                        "      --precision M            0 complex128 per gate, 1 complex64 fused, 2 complex128 fused\n"
                        "      --block-qubits K         qubits per fused block (precision 1, 2)\n"
                        "  -h, --help\n");
    }
} // namespace

int main(int argc, char **argv)
{
    try
    {
        int num_samples = 1;
        bool reversed = false, have_mapping = false, have_seed = false;
        std::string mapping_text, output_path;
        unsigned long long seed = 0;
        int precision = 0, block_qubits = 5;
        static const option opts[] = {
            {"num-samples", required_argument, nullptr, 's'},
            {"mapping", required_argument, nullptr, 'm'},
            {"reversed-mapping", no_argument, nullptr, 'r'},
            {"output-statevector", required_argument, nullptr, 'o'},
            {"seed", required_argument, nullptr, 'S'},
            {"precision", required_argument, nullptr, 'p'},
            {"block-qubits", required_argument, nullptr, 'k'},
            {"help", no_argument, nullptr, 'h'},
            {nullptr, 0, nullptr, 0}};
        int c;
        while ((c = getopt_long(argc, argv, "s:rh", opts, nullptr)) != -1)
        {
            switch (c)
            {
            case 's':
                num_samples = std::stoi(optarg);
                break;
            case 'm':
                mapping_text = optarg;
                have_mapping = true;
                break;
            case 'r':
                reversed = true;
                break;
            case 'o':
                output_path = optarg;
                break;
            case 'S':
                seed = std::stoull(optarg);
                have_seed = true;
                break;
            case 'p':
                precision = std::stoi(optarg);
                break;
            case 'k':
                block_qubits = std::stoi(optarg);
                break;
            case 'h':
                usage(stdout);
                return 0;
            default:
                usage(stderr);
                return 2;
            }
        }
        if (argc - optind != 1)
            throw std::runtime_error("user circuit shared object path is required");
        if (num_samples <= 0)
            throw std::runtime_error("num_samples must be greater than 0");
        qcs_set_exception_callback(on_error);

        qcs_simulator *raw = nullptr;
        must(qcs_simulator_create(&raw), "qcs_simulator_create");
        std::unique_ptr<qcs_simulator, int (*)(qcs_simulator *)> sim(raw, qcs_simulator_destroy);
        if (have_seed)
            must(qcs_ext_set_seed(sim.get(), seed), "qcs_ext_set_seed");
        if (precision != 0)
            must(qcs_ext_set_precision(sim.get(), precision, block_qubits), "qcs_ext_set_precision");

        char *abs = realpath(argv[optind], nullptr);
        if (abs == nullptr)
            throw std::runtime_error("realpath returned NULL");
        std::unique_ptr<char, void (*)(void *)> abs_guard(abs, free);
        void *dl = dlopen(abs, RTLD_LAZY);
        if (dl == nullptr)
            throw std::runtime_error(std::string("dlopen failed: ") + dlerror());
        std::unique_ptr<void, int (*)(void *)> dl_guard(dl, dlclose);
        auto circuit_init = reinterpret_cast<void (*)(qcs_simulator *)>(dlsym(dl, "circuit_init"));
        auto circuit_run = reinterpret_cast<void (*)(qcs_simulator *)>(dlsym(dl, "circuit_run"));
        if (!circuit_init || !circuit_run)
            throw std::runtime_error("the plug-in must export circuit_init and circuit_run");

        circuit_init(sim.get());

        bit_t nq = 0;
        must(qcs_simulator_get_num_qubits(sim.get(), &nq), "qcs_simulator_get_num_qubits");
        std::vector<bit_num_t> mapping;
        if (have_mapping)
            mapping = parse_mapping(mapping_text);
        else
            for (int q = 0; q < nq; ++q)
                mapping.push_back((bit_num_t)q);
        if (reversed)
            mapping = invert(mapping, nq);
        must(qcs_simulator_set_mapping(sim.get(), mapping.data(), (bit_num_t)mapping.size()), "qcs_simulator_set_mapping");
        must(qcs_simulator_allocate_memory(sim.get()), "qcs_simulator_allocate_memory");

        bit_t e1 = 0, e2 = 0;
        must(qcs_simulator_event_create(sim.get(), &e1), "qcs_simulator_event_create");
        must(qcs_simulator_event_create(sim.get(), &e2), "qcs_simulator_event_create");
        for (int s = 0; s < num_samples; ++s)
        {
            must(qcs_simulator_event_record(sim.get(), e1), "qcs_simulator_event_record");
            circuit_run(sim.get());
            must(qcs_simulator_event_record(sim.get(), e2), "qcs_simulator_event_record");
            double elapsed = 0;
            must(qcs_simulator_event_get_elapsed_time(sim.get(), e1, e2, &elapsed), "qcs_simulator_event_get_elapsed_time");
            bit_t nc = 0;
            must(qcs_simulator_get_num_clbits(sim.get(), &nc), "qcs_simulator_get_num_clbits");
            std::vector<char> bits(nc + 1);
            must(qcs_simulator_get_clbits_string(sim.get(), bits.data()), "qcs_simulator_get_clbits_string");
            std::printf("{\"sample_num\": %d, \"clbits\": \"%s\", \"elapsed_time\": %.18g}\n", s, bits.data(), elapsed);
            std::fflush(stdout);
            if (s != num_samples - 1)
            {
                must(qcs_simulator_reinitialize_mapping(sim.get()), "qcs_simulator_reinitialize_mapping");
                must(qcs_simulator_set_zero_state(sim.get()), "qcs_simulator_set_zero_state");
                must(qcs_simulator_reset_clbits(sim.get()), "qcs_simulator_reset_clbits");
                must(qcs_simulator_reset_measurement_state(sim.get()), "qcs_simulator_reset_measurement_state");
            }
        }
        if (!output_path.empty())
            must(qcs_simulator_save_statevector(sim.get(), output_path.c_str()), "qcs_simulator_save_statevector");
        return 0;
    }
    catch (const std::exception &e)
    {
        std::fprintf(stderr, "qcs: %s\n", e.what());
        return 1;
    }
}
