// io.hpp -- the reference's on-disk formats and logging layer, over the particle classes:
//   io_utils      include/io/io_utils.h:19-190, src/io/io_utils.cu:31-263
//   log groups    include/io/base_log_groups.h:12-172, src/io/base_log_groups.cu:7-150
//   EnergyLog     src/io/energy_log.cu:5-59 (<root>/system/energy.csv)
//   ConsoleLog    src/io/console_log.cu:3-42
//   StateLog      src/io/state_log.cu:5-65 (<root>/trajectories/t<step>/, <root>/system/restart/)
//   Orchestrator  include/io/orchestrator.h:15-65, src/io/orchestrator.cu:34-79
//   IOManager     include/io/io_manager.h, src/io/io_manager.cu:7-199
//   load_configs  include/io/arg_parsing.h:7-63
//
// Formats (as the reference writes them):
//   <name>.dat   one row per element in the static particle (or vertex) order; doubles in
//                std::scientific with 17 digits after the point (printf "%.17e"), longs in
//                decimal; 2-D arrays as "x<TAB>y"; every row ends in '\n'.
//   step.dat     the step number, no newline (restart directories only).
//   energy.csv   a header of the log names joined by ',' then one row per logged step,
//                each value "%.17e" (the step too), joined by ','.
//   *.json       ConfigDict::save (4-space indent, sorted keys).
// A name "X/M" logs X divided by M = N (particles), Nv (vertices) or dof.
//
// Errors: where the reference prints a message and exits, these throw
// std::runtime_error carrying the same message.
#pragma once
#include <condition_variable>
#include <deque>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "array_data.hpp"
#include "config_dict.hpp"
#include "integrator.hpp"
#include "particle.hpp"

constexpr long DECIMAL_PRECISION = 17;  // include/constants.h

// ---------------------------------------------------------------- io_utils
std::ifstream open_input_file(std::string file_name);
std::ofstream open_output_file(std::string file_name, bool overwrite = false);
void make_dir(const std::string& dir_name, bool overwrite = false);
std::filesystem::path get_path(std::filesystem::path root_path, std::string name);
bool contains_substrings(const std::string& string, const std::vector<std::string>& substrings);
// largest N among the entries <prefix>N of a directory
long get_largest_file_index(std::string dir_name, std::string file_prefix = "");
// the entry <prefix><file_index> of trajectory_root, or the largest when file_index == -1
std::tuple<std::filesystem::path, long> get_trajectory_frame_path(
    std::filesystem::path trajectory_root, std::string file_prefix, long file_index);
void write_json_to_file(const std::string& file_name, const ConfigValue& data);

// scatter: new[index[i]] = data[i] (io_utils.cu:174-231)
void reorder_array(ArrayData& array_data, const ArrayData& reorder_index_data);

template <typename T>
void write_1d_array_to_file(std::ofstream& out, const std::vector<T>& data, int precision) {
    for (const auto& v : data) out << std::scientific << std::setprecision(precision) << v << "\n";
}
template <typename T>
void write_2d_array_to_file(std::ofstream& out,
                            const std::pair<std::vector<T>, std::vector<T>>& data,
                            int precision) {
    if (data.first.size() != data.second.size())
        throw std::runtime_error("Error: 2D data vectors have mismatched sizes.");
    for (size_t i = 0; i < data.first.size(); ++i)
        out << std::scientific << std::setprecision(precision) << data.first[i] << "\t"
            << data.second[i] << "\n";
}
void write_array_data_to_file(const std::string& file_name, ArrayData& array_data,
                              long precision);

// num_rows * num_cols whitespace-separated values, row-major (io_utils.h:116-137)
template <typename T>
std::vector<T> read_array_from_file(const std::string& file_name, long num_rows, long num_cols) {
    std::vector<T> data((size_t)(num_rows * num_cols));
    std::ifstream in(file_name);
    if (!in.is_open())
        throw std::runtime_error("read_array_from_file: Error: could not open input file " +
                                 file_name);
    for (long i = 0; i < num_rows * num_cols; ++i)
        if (!(in >> data[(size_t)i]))
            throw std::runtime_error("Error: insufficient data in file " + file_name);
    return data;
}

// ---------------------------------------------------------------- log group configs
ConfigDict get_log_group_config_dict();
ConfigDict config_from_names_lin(std::vector<std::string> log_names, long num_steps,
                                 long num_saves, std::string group_name);
ConfigDict config_from_names_log(std::vector<std::string> log_names, long num_steps,
                                 long num_saves, long min_save_decade, std::string group_name);
ConfigDict config_from_names_lin_everyN(std::vector<std::string> log_names, long save_freq,
                                        std::string group_name);
ConfigDict config_from_names(std::vector<std::string> log_names, std::string group_name);

// ---------------------------------------------------------------- orchestrator
class Orchestrator {
   public:
    Orchestrator(Particle& particle, Integrator* integrator = nullptr);
    void reset_dependency_status();
    void handle_dependencies(std::string log_name);
    bool is_dependent(std::string log_name);
    double apply_modifier(std::string& modifier, double value);
    // step, KE, PE, TE, T, phi, Zp, Zv, P (anything else: a message and 0.0)
    double get_value(const std::string& unmodified_log_name, long step);
    // arrays arrive from the particle already in the static order
    bool arrays_need_reordering = false;
    std::unordered_map<std::string, ArrayData> get_reorder_index_data();
    ArrayData get_array_data(const std::string& unmodified_log_name);

    Particle& particle;
    Integrator* integrator;
};

// ---------------------------------------------------------------- log groups
class BaseLogGroup {
   public:
    BaseLogGroup(ConfigDict config, Orchestrator& orchestrator);
    virtual ~BaseLogGroup() = default;
    // "lin": every save_freq steps; "log": save_freq grows by `decade` each decade and
    // restarts from min_save_decade every reset_save_decade steps (base_log_groups.cu:80-102)
    void update_log_status(long step);
    virtual void define_dependencies();
    void handle_dependencies();
    virtual void gather_data(long step) = 0;
    virtual void log(long step) = 0;
    virtual std::unique_ptr<BaseLogGroup> snapshot() const = 0;

    ConfigDict config;
    Orchestrator& orchestrator;
    std::vector<std::string> log_names;
    std::string save_style;
    long save_freq, reset_save_decade, min_save_decade, multiple, decade;
    std::string group_name;
    bool should_log = false;
    bool parallel = false;
    bool has_dependencies = false;
    std::set<std::string> dependencies;
};

class ScalarLog : public BaseLogGroup {
   public:
    ScalarLog(ConfigDict config, Orchestrator& orchestrator);
    void define_dependencies() override;
    void gather_data(long step) override;
    bool is_modified(std::string log_name);
    std::string get_modifier(std::string log_name);
    std::vector<std::string> get_unmodified_log_names();

    std::vector<std::string> unmodified_log_names;
    std::unordered_map<std::string, double> gathered_data;
    std::string modifier = "/";
};

class EnergyLog : public ScalarLog {
   public:
    EnergyLog(ConfigDict config, Orchestrator& orchestrator, const std::string& file_name,
              bool overwrite = false);
    void write_header();
    void log(long step) override;
    std::unique_ptr<BaseLogGroup> snapshot() const override;

    std::string file_name;
    std::ofstream log_file;
    std::string delimiter = ",";
    long precision = DECIMAL_PRECISION;
    bool has_header = false;
};

class ConsoleLog : public ScalarLog {
   public:
    ConsoleLog(ConfigDict config, Orchestrator& orchestrator);
    void write_header();
    void log(long step) override;
    std::unique_ptr<BaseLogGroup> snapshot() const override;

    std::string delimiter = "|";
    int precision = 3;
    int width = 10;
    int header_log_freq = 10;
    int last_header_log = 11;
};

class StateLog : public BaseLogGroup {
   public:
    StateLog(ConfigDict config, Orchestrator& orchestrator, const std::string& root,
             const std::string& indexed_file_prefix, const std::string& extension,
             bool is_restart = false);
    void gather_data(long step) override;
    void write_values(const std::filesystem::path& root, long step);
    void log(long step) override;
    void write_state_to_path(const std::filesystem::path& path);
    std::unique_ptr<BaseLogGroup> snapshot() const override;

    long precision = DECIMAL_PRECISION;
    std::string root, indexed_file_prefix, extension;
    std::unordered_map<std::string, ArrayData> gathered_data;
    std::unordered_map<std::string, ArrayData> reorder_index_data;
    bool is_restart;

   private:
    // shared by a restart group and its queued snapshots: restart writes into the one
    // restart directory one at a time, and a write older than the last one is dropped
    struct RestartOrder {
        std::mutex mu;
        long last = std::numeric_limits<long>::min();
    };
    std::shared_ptr<RestartOrder> order_ = std::make_shared<RestartOrder>();
};

// a fixed set of worker threads; shutdown() (also the destructor) runs every queued task
class ThreadPool {
   public:
    explicit ThreadPool(long num_threads);
    ~ThreadPool();
    void enqueue(std::function<void()> task);
    void shutdown();

   private:
    std::vector<std::thread> workers_;
    std::deque<std::function<void()>> tasks_;
    std::mutex mu_;
    std::condition_variable cv_;
    bool stop_ = false;
};

// ---------------------------------------------------------------- IOManager
// Log groups by "group_name": "energy" (<root>/system/energy.csv), "console", "state"
// (<root>/trajectories/t<step>/), "restart" (<root>/system/restart/; with overwrite the
// state at construction also goes to <root>/system/init/).  overwrite removes <root>
// first.  State logs are written by num_threads workers when num_threads > 1 (the data
// is gathered on the calling thread first).
class IOManager {
   public:
    IOManager(std::vector<ConfigDict> log_configs, Particle& particle,
              Integrator* integrator = nullptr, std::string root = "", long num_threads = 1,
              bool overwrite = true);
    ~IOManager();
    // IOManager::log (io_manager.cu:118-159) = prepare + write
    void log(long step, bool force = false);
    // advances every group's schedule to `step`; true when something is to be logged
    bool prepare(long step, bool force = false);
    // gathers and writes what prepare(step, force) selected
    void write(long step, bool force = false);
    void write_restart_file(long step, std::string dir_name);
    void write_log_configs(std::filesystem::path path);
    void write_particle_config(std::filesystem::path path);
    void write_integrator_config(std::filesystem::path path);
    void write_params();

    std::filesystem::path system_dir_path, trajectory_dir_path, energy_file_path;
    std::filesystem::path init_dir_path, restart_dir_path;

   private:
    void init_path(std::filesystem::path& path, const std::string& path_name);

    Particle& particle;
    Integrator* integrator;
    Orchestrator orchestrator;
    std::vector<ConfigDict> log_configs;
    bool overwrite;
    long num_threads;
    bool use_parallel;
    std::string root;
    std::filesystem::path root_path;
    std::vector<BaseLogGroup*> log_groups;
    StateLog* state_log = nullptr;
    ThreadPool thread_pool;

    std::string energy_file_extension = ".csv";
    std::string state_file_extension = ".dat";
    std::string indexed_file_prefix = "t";
    std::string energy_file_name = "energy";
    std::string system_dir_name = "system";
    std::string trajectory_dir_name = "trajectories";
    std::string restart_dir_name = "restart";
    std::string init_dir_name = "init";
};

// ---------------------------------------------------------------- load_configs
// argv[1] = a run directory holding system/run_config.json with "input_dir" and
// "output_dir": only output_dir -> a new particle from output_dir/system/
// particle_config.json (step 0); both -> the particle loaded from input_dir (restart
// state, or init when use_restart is false), step 0; only input_dir -> resume from
// input_dir at its restart step.  The console / energy / state log configs come from
// <config root>/system/*_log_config.json.  Returns (particle, step, console, energy,
// state, run config).
std::tuple<std::unique_ptr<Particle>, long, ConfigDict, ConfigDict, ConfigDict, ConfigDict>
load_configs(int argc, char** argv, bool use_restart = true);
