// io.cpp -- on-disk formats, log groups and IOManager (see io.hpp for the reference map).
#include "io.hpp"

#include <algorithm>
#include <sstream>
#include <stdexcept>

#include "factory.hpp"

// ---------------------------------------------------------------- io_utils (io_utils.cu:31-165)
std::ifstream open_input_file(std::string file_name) {
    std::ifstream in(file_name);
    if (!in.is_open())
        throw std::runtime_error("ERROR: open_input_file: could not open: " + file_name);
    return in;
}

std::ofstream open_output_file(std::string file_name, bool overwrite) {
    std::ofstream out(file_name, overwrite ? (std::ios::out | std::ios::trunc)
                                           : (std::ios::out | std::ios::app));
    if (!out.is_open())
        throw std::runtime_error("ERROR: open_output_file: could not open: " + file_name);
    return out;
}

// an existing directory is kept as it is, with or without overwrite (io_utils.cu:54-69)
void make_dir(const std::string& dir_name, bool overwrite) {
    (void)overwrite;
    if (std::filesystem::exists(dir_name)) return;
    std::error_code ec;
    std::filesystem::create_directories(dir_name, ec);
    if (ec)
        throw std::runtime_error("ERROR: make_dir: issue creating directory: " + dir_name +
                                 " Error code: " + ec.message());
}

std::filesystem::path get_path(std::filesystem::path root_path, std::string name) {
    std::filesystem::path p = root_path / name;
    if (!std::filesystem::exists(p))
        throw std::runtime_error("ERROR: get_path: path does not exist: " + p.string());
    return p;
}

bool contains_substrings(const std::string& s, const std::vector<std::string>& subs) {
    for (const auto& x : subs)
        if (s.find(x) != std::string::npos) return true;
    return false;
}

long get_largest_file_index(std::string dir_name, std::string file_prefix) {
    long largest = -1;
    for (const auto& entry : std::filesystem::directory_iterator(dir_name)) {
        std::string f = entry.path().filename().string();
        if (f.find(file_prefix) == 0) largest = std::max(largest, std::stol(f.substr(file_prefix.size())));
    }
    if (largest == -1)
        throw std::runtime_error("ERROR: get_largest_file_index: could not find any indexed files in " +
                                 dir_name + " with matching prefix '" + file_prefix + "'");
    return largest;
}

std::tuple<std::filesystem::path, long> get_trajectory_frame_path(
    std::filesystem::path trajectory_root, std::string file_prefix, long file_index) {
    std::filesystem::path best_path;
    long best_frame = -1;
    for (const auto& entry : std::filesystem::directory_iterator(trajectory_root)) {
        const std::string f = entry.path().filename().string();
        if (f.rfind(file_prefix, 0) != 0) continue;
        const long frame = std::stol(f.substr(file_prefix.size()));
        if (file_index == -1) {
            if (frame > best_frame) {
                best_frame = frame;
                best_path = entry.path();
            }
        } else if (frame == file_index) {
            return std::make_tuple(entry.path(), frame);
        }
    }
    if (file_index == -1 && !best_path.empty()) return std::make_tuple(best_path, best_frame);
    throw std::runtime_error("ERROR: get_trajectory_frame_path: no matching frame path for prefix '" +
                             file_prefix + "' and index " + std::to_string(file_index) + " in " +
                             trajectory_root.string());
}

void write_json_to_file(const std::string& file_name, const ConfigValue& data) {
    std::ofstream out = open_output_file(file_name, true);
    out << data.dump(4);
}

namespace {
template <typename T>
void scatter(std::vector<T>& data, const std::vector<long>& index) {
    std::vector<T> out(data.size());
    for (size_t i = 0; i < data.size(); ++i) out[(size_t)index[i]] = data[i];
    data.swap(out);
}
}  // namespace

void reorder_array(ArrayData& a, const ArrayData& index_data) {
    const auto& index = std::get<std::vector<long>>(index_data.data);
    if (a.size[1] == 1) {
        if (a.type == DataType::Double) scatter(std::get<std::vector<double>>(a.data), index);
        else scatter(std::get<std::vector<long>>(a.data), index);
    } else if (a.size[1] == 2) {
        if (a.type == DataType::Double) {
            auto& d = std::get<std::pair<std::vector<double>, std::vector<double>>>(a.data);
            scatter(d.first, index);
            scatter(d.second, index);
        } else {
            auto& d = std::get<std::pair<std::vector<long>, std::vector<long>>>(a.data);
            scatter(d.first, index);
            scatter(d.second, index);
        }
    } else {
        throw std::runtime_error(
            "ERROR: reorder_array: array_data has invalid number of dimensions: " +
            std::to_string(a.size[1]));
    }
}

void write_array_data_to_file(const std::string& file_name, ArrayData& a, long precision) {
    std::ofstream out(file_name);
    if (!out.is_open())
        throw std::runtime_error("write_array_data_to_file: Error: could not open output file " +
                                 file_name);
    const int p = (int)precision;
    if (a.size[1] == 1) {
        if (a.type == DataType::Double) write_1d_array_to_file(out, std::get<std::vector<double>>(a.data), p);
        else write_1d_array_to_file(out, std::get<std::vector<long>>(a.data), p);
    } else if (a.size[1] == 2) {
        if (a.type == DataType::Double)
            write_2d_array_to_file(out, std::get<std::pair<std::vector<double>, std::vector<double>>>(a.data), p);
        else
            write_2d_array_to_file(out, std::get<std::pair<std::vector<long>, std::vector<long>>>(a.data), p);
    } else {
        throw std::runtime_error(
            "write_array_data_to_file: Error: Unsupported array dimensionality: " +
            std::to_string(a.size[1]));
    }
}

// ---------------------------------------------------------------- log group configs
// (base_log_groups.h:12-24, base_log_groups.cu:7-46)
ConfigDict get_log_group_config_dict() {
    ConfigDict c;
    c["log_names"] = std::vector<std::string>();
    c["save_style"] = "lin";
    c["save_freq"] = 1;
    c["reset_save_decade"] = 10;
    c["min_save_decade"] = 10;
    c["multiple"] = 0;
    c["group_name"] = "none";
    c["decade"] = 10;
    return c;
}

ConfigDict config_from_names_lin(std::vector<std::string> names, long num_steps, long num_saves,
                                 std::string group_name) {
    ConfigDict c = get_log_group_config_dict();
    c["log_names"] = names;
    c["save_style"] = "lin";
    c["save_freq"] = static_cast<long>(num_steps / num_saves);
    if (c.at("save_freq").get<long>() == 0) c["save_freq"] = 1;
    c["group_name"] = group_name;
    return c;
}

ConfigDict config_from_names_log(std::vector<std::string> names, long num_steps, long num_saves,
                                 long min_save_decade, std::string group_name) {
    ConfigDict c = get_log_group_config_dict();
    c["log_names"] = names;
    c["save_style"] = "log";
    c["reset_save_decade"] = static_cast<long>(num_steps / num_saves);
    if (c.at("reset_save_decade").get<long>() == 0) c["reset_save_decade"] = 1;
    c["min_save_decade"] = min_save_decade;
    c["group_name"] = group_name;
    return c;
}

ConfigDict config_from_names_lin_everyN(std::vector<std::string> names, long save_freq,
                                        std::string group_name) {
    ConfigDict c = get_log_group_config_dict();
    c["log_names"] = names;
    c["save_style"] = "lin";
    c["save_freq"] = save_freq;
    c["group_name"] = group_name;
    return c;
}

ConfigDict config_from_names(std::vector<std::string> names, std::string group_name) {
    ConfigDict c = get_log_group_config_dict();
    c["log_names"] = names;
    c["group_name"] = group_name;
    return c;
}

// ---------------------------------------------------------------- Orchestrator
// (orchestrator.cu:34-79, orchestrator.h:36-58)
Orchestrator::Orchestrator(Particle& p, Integrator* i) : particle(p), integrator(i) {}

void Orchestrator::reset_dependency_status() { particle.reset_dependency_status(); }
void Orchestrator::handle_dependencies(std::string log_name) {
    particle.calculate_dependencies(log_name);
}
bool Orchestrator::is_dependent(std::string log_name) { return particle.is_dependent(log_name); }

double Orchestrator::apply_modifier(std::string& modifier, double value) {
    if (modifier == "N") return value / particle.n_particles;
    if (modifier == "Nv") return value / particle.n_vertices;
    if (modifier == "dof") return value / particle.n_dof;
    std::cerr << "Orchestrator::apply_modifier: Modifier not recognized: " << modifier << std::endl;
    return value;
}

double Orchestrator::get_value(const std::string& name, long step) {
    if (name == "step") return (double)step;
    if (name == "KE") return particle.totalKineticEnergy();
    if (name == "PE") return particle.totalPotentialEnergy();
    if (name == "TE") return particle.totalEnergy();
    if (name == "T") return particle.calculateTemperature();
    if (name == "phi") return particle.getPackingFraction();
    if (name == "Zp") return particle.getContactCount();
    if (name == "Zv") return particle.getVertexContactCount();
    if (name == "P") return particle.getPressure();
    std::cerr << "Orchestrator::get_value: Log name not recognized: " << name << std::endl;
    return 0.0;
}

std::unordered_map<std::string, ArrayData> Orchestrator::get_reorder_index_data() {
    std::unordered_map<std::string, ArrayData> out;
    for (const std::string& n : particle.get_reorder_arrays()) out[n] = particle.getArrayData(n);
    return out;
}

ArrayData Orchestrator::get_array_data(const std::string& name) {
    return particle.getArrayData(name);
}

// ---------------------------------------------------------------- log groups
BaseLogGroup::BaseLogGroup(ConfigDict c, Orchestrator& o) : config(c), orchestrator(o) {
    log_names = config.at("log_names").get<std::vector<std::string>>();
    save_style = config.at("save_style").get<std::string>();
    save_freq = config.at("save_freq").get<long>();
    reset_save_decade = config.at("reset_save_decade").get<long>();
    min_save_decade = config.at("min_save_decade").get<long>();
    multiple = config.at("multiple").get<long>();
    decade = config.at("decade").get<long>();
    group_name = config.at("group_name").get<std::string>();
}

void BaseLogGroup::define_dependencies() {
    for (const std::string& n : log_names)
        if (orchestrator.is_dependent(n)) {
            dependencies.insert(n);
            has_dependencies = true;
        }
}

void BaseLogGroup::handle_dependencies() {
    for (const std::string& n : dependencies) orchestrator.handle_dependencies(n);
}

// BaseLogGroup::update_log_status (base_log_groups.cu:80-102)
void BaseLogGroup::update_log_status(long step) {
    if (save_style == "lin") {
        should_log = step % save_freq == 0;
    } else if (save_style == "log") {
        if (step > multiple * reset_save_decade) {
            save_freq = min_save_decade;
            multiple += 1;
        }
        if ((step - (multiple - 1) * reset_save_decade) > save_freq * decade) save_freq *= decade;
        should_log = (step - (multiple - 1) * reset_save_decade) % save_freq == 0;
    } else {
        throw std::runtime_error("ERROR: LogManager::update_log_status: Invalid save style: " +
                                 save_style);
    }
}

ScalarLog::ScalarLog(ConfigDict c, Orchestrator& o) : BaseLogGroup(c, o) {
    unmodified_log_names = get_unmodified_log_names();
}

void ScalarLog::define_dependencies() {
    for (const std::string& n : unmodified_log_names)
        if (orchestrator.is_dependent(n)) {
            dependencies.insert(n);
            has_dependencies = true;
        }
}

bool ScalarLog::is_modified(std::string n) { return n.find(modifier) != std::string::npos; }

std::string ScalarLog::get_modifier(std::string n) {
    size_t pos = n.find(modifier);
    return pos == std::string::npos ? "" : n.substr(pos + 1);
}

std::vector<std::string> ScalarLog::get_unmodified_log_names() {
    std::vector<std::string> out;
    for (const auto& n : log_names) out.push_back(n.substr(0, n.find(modifier)));
    return out;
}

void ScalarLog::gather_data(long step) {
    for (size_t i = 0; i < unmodified_log_names.size(); ++i) {
        double v = orchestrator.get_value(unmodified_log_names[i], step);
        if (is_modified(log_names[i])) {
            std::string mod = get_modifier(log_names[i]);
            v = orchestrator.apply_modifier(mod, v);
        }
        gathered_data[log_names[i]] = v;
    }
}

// EnergyLog (energy_log.cu:5-59).  The header goes into a new or empty file only.
EnergyLog::EnergyLog(ConfigDict c, Orchestrator& o, const std::string& fname, bool overwrite)
    : ScalarLog(c, o), file_name(fname) {
    std::error_code ec;
    bool had_content = !overwrite && std::filesystem::exists(fname) &&
                       std::filesystem::file_size(fname, ec) > 0;
    log_file = open_output_file(fname, overwrite);
    if (had_content) has_header = true;
    else write_header();
}

void EnergyLog::write_header() {
    for (size_t i = 0; i < log_names.size(); ++i) {
        log_file << log_names[i];
        if (i + 1 < log_names.size()) log_file << delimiter;
    }
    log_file << "\n";
    log_file.flush();
    has_header = true;
}

void EnergyLog::log(long step) {
    (void)step;
    for (size_t i = 0; i < log_names.size(); ++i) {
        log_file << std::scientific << std::setprecision((int)precision)
                 << gathered_data[log_names[i]];
        if (i + 1 < log_names.size()) log_file << delimiter;
    }
    log_file << "\n";
    log_file.flush();
    gathered_data.clear();
}

std::unique_ptr<BaseLogGroup> EnergyLog::snapshot() const {
    auto s = std::make_unique<EnergyLog>(config, orchestrator, file_name, false);
    s->gathered_data = gathered_data;
    s->has_header = has_header;
    return s;
}

// ConsoleLog (console_log.cu:3-42)
ConsoleLog::ConsoleLog(ConfigDict c, Orchestrator& o) : ScalarLog(c, o) {}

void ConsoleLog::write_header() {
    std::ostringstream out;
    std::string rule((size_t)(width * log_names.size() + (log_names.size() - 1)), '_');
    out << rule << std::endl;
    for (size_t i = 0; i < log_names.size(); ++i) {
        out << std::setw(width) << log_names[i];
        if (i + 1 < log_names.size()) out << delimiter;
    }
    out << std::endl << rule << std::endl;
    std::cout << out.str();
}

void ConsoleLog::log(long step) {
    (void)step;
    if (last_header_log > header_log_freq) {
        write_header();
        last_header_log = 0;
    }
    last_header_log += 1;
    std::ostringstream out;
    for (size_t i = 0; i < log_names.size(); ++i) {
        out << std::setw(width) << std::scientific << std::setprecision(precision)
            << gathered_data[log_names[i]];
        if (i + 1 < log_names.size()) out << delimiter;
    }
    std::cout << out.str() << std::endl;
    gathered_data.clear();
}

std::unique_ptr<BaseLogGroup> ConsoleLog::snapshot() const {
    return std::make_unique<ConsoleLog>(*this);
}

// StateLog (state_log.cu:5-65)
StateLog::StateLog(ConfigDict c, Orchestrator& o, const std::string& r, const std::string& prefix,
                   const std::string& ext, bool restart)
    : BaseLogGroup(c, o), root(r), indexed_file_prefix(prefix), extension(ext), is_restart(restart) {
    parallel = true;
}

void StateLog::gather_data(long step) {
    (void)step;
    if (orchestrator.arrays_need_reordering) reorder_index_data = orchestrator.get_reorder_index_data();
    for (const auto& n : log_names) {
        ArrayData a = orchestrator.get_array_data(n);
        if (a.name != "NULL") gathered_data[n] = std::move(a);
    }
}

void StateLog::write_values(const std::filesystem::path& dir, long step) {
    for (auto& [name, a] : gathered_data) {
        std::filesystem::path file = dir / (name + extension);
        if (orchestrator.arrays_need_reordering &&
            reorder_index_data.find(a.index_array_name) != reorder_index_data.end())
            reorder_array(a, reorder_index_data[a.index_array_name]);
        write_array_data_to_file(file.string(), a, precision);
    }
    if (is_restart) {
        std::ofstream step_file((dir / ("step" + extension)).string());
        step_file << step;
    }
}

void StateLog::log(long step) {
    std::filesystem::path dir = is_restart
                                    ? std::filesystem::path(root) / "restart"
                                    : std::filesystem::path(root) /
                                          (indexed_file_prefix + std::to_string(step));
    if (is_restart) {
        // queued restart writes may run on several workers: one at a time, newest wins
        std::lock_guard<std::mutex> g(order_->mu);
        if (step >= order_->last) {
            make_dir(dir.string(), true);
            write_values(dir, step);
            order_->last = step;
        }
    } else {
        make_dir(dir.string(), true);
        write_values(dir, step);
    }
    gathered_data.clear();
    reorder_index_data.clear();
}

void StateLog::write_state_to_path(const std::filesystem::path& path) { write_values(path, 0); }

std::unique_ptr<BaseLogGroup> StateLog::snapshot() const { return std::make_unique<StateLog>(*this); }

// ---------------------------------------------------------------- ThreadPool
ThreadPool::ThreadPool(long num_threads) {
    for (long i = 0; i < num_threads; ++i)
        workers_.emplace_back([this] {
            while (true) {
                std::function<void()> task;
                {
                    std::unique_lock<std::mutex> lk(mu_);
                    cv_.wait(lk, [this] { return stop_ || !tasks_.empty(); });
                    if (tasks_.empty()) return;  // stop_ and drained
                    task = std::move(tasks_.front());
                    tasks_.pop_front();
                }
                try {
                    task();
                } catch (const std::exception& e) {  // a failed write must not end the process
                    std::cerr << "ERROR: IOManager worker: " << e.what() << std::endl;
                }
            }
        });
}

ThreadPool::~ThreadPool() { shutdown(); }

void ThreadPool::enqueue(std::function<void()> task) {
    if (workers_.empty()) {
        task();
        return;
    }
    {
        std::lock_guard<std::mutex> lk(mu_);
        tasks_.push_back(std::move(task));
    }
    cv_.notify_one();
}

void ThreadPool::shutdown() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_) return;
        stop_ = true;
    }
    cv_.notify_all();
    for (auto& w : workers_)
        if (w.joinable()) w.join();
}

// ---------------------------------------------------------------- IOManager (io_manager.cu:7-199)
IOManager::IOManager(std::vector<ConfigDict> configs, Particle& p, Integrator* integ,
                     std::string root_dir, long threads, bool ow)
    : particle(p),
      integrator(integ),
      orchestrator(p, integ),
      log_configs(configs),
      overwrite(ow),
      num_threads(threads),
      use_parallel(threads > 1),
      root(root_dir),
      thread_pool(threads > 1 ? threads : 0) {
    root_path = std::filesystem::path(root);
    if (!root_path.empty()) {
        if (overwrite) {
            std::filesystem::remove_all(root_path);
            std::filesystem::create_directories(root_path);
        } else if (std::filesystem::exists(root_path)) {
            std::cout << "ERROR: IOManager::IOManager: root path " << root_path
                      << " already exists and overwriting is disabled!\n";
        }
    }
    for (auto& config : log_configs) {
        const std::string group = config.at("group_name").get<std::string>();
        if (group == "energy") {
            if (system_dir_path.empty()) {
                init_path(system_dir_path, system_dir_name);
                make_dir(system_dir_path.string(), overwrite);
            }
            energy_file_path = system_dir_path / (energy_file_name + energy_file_extension);
            log_groups.push_back(new EnergyLog(config, orchestrator, energy_file_path.string(), overwrite));
        } else if (group == "console") {
            log_groups.push_back(new ConsoleLog(config, orchestrator));
        } else if (group == "state") {
            if (trajectory_dir_path.empty()) {
                init_path(trajectory_dir_path, trajectory_dir_name);
                make_dir(trajectory_dir_path.string(), overwrite);
            }
            log_groups.push_back(new StateLog(config, orchestrator, trajectory_dir_path.string(),
                                              indexed_file_prefix, state_file_extension));
        } else if (group == "restart") {
            if (system_dir_path.empty()) {
                init_path(system_dir_path, system_dir_name);
                make_dir(system_dir_path.string(), overwrite);
            }
            state_log = new StateLog(config, orchestrator, system_dir_path.string(), "",
                                     state_file_extension, true);
            log_groups.push_back(state_log);
        } else {
            std::cerr << "ERROR: IOManager::IOManager: " << group
                      << " is not a valid log group name\n";
        }
    }
    for (auto* g : log_groups) g->define_dependencies();
    if (state_log != nullptr && overwrite) {
        init_dir_path = system_dir_path / init_dir_name;
        make_dir(init_dir_path.string(), true);
        restart_dir_path = system_dir_path / restart_dir_name;
        make_dir(restart_dir_path.string(), true);
        orchestrator.reset_dependency_status();
        state_log->handle_dependencies();
        state_log->gather_data(0);
        state_log->write_state_to_path(init_dir_path);
        state_log->gathered_data.clear();
    }
}

IOManager::~IOManager() {
    thread_pool.shutdown();
    for (auto* g : log_groups) delete g;
}

void IOManager::init_path(std::filesystem::path& path, const std::string& name) {
    if (root.empty()) {
        std::cerr << "ERROR: IOManager::init_path: root is empty. path_name=" << name << std::endl;
        return;
    }
    path = std::filesystem::path(root) / name;
}

void IOManager::log(long step, bool force) {
    if (prepare(step, force)) write(step, force);
}

bool IOManager::prepare(long step, bool force) {
    bool required = false;
    for (auto* g : log_groups) {
        g->update_log_status(step);
        if (g->should_log || force) required = true;
    }
    return required;
}

void IOManager::write(long step, bool force) {
    orchestrator.reset_dependency_status();
    for (auto* g : log_groups)
        if (g->has_dependencies) g->handle_dependencies();
    for (auto* g : log_groups)
        if (g->should_log || force) g->gather_data(step);
    for (auto* g : log_groups) {
        if (!(g->should_log || force)) continue;
        if (g->parallel && use_parallel) {
            std::shared_ptr<BaseLogGroup> snap(g->snapshot());
            thread_pool.enqueue([snap, step]() { snap->log(step); });
            // the snapshot owns the gathered arrays now
            if (auto* s = dynamic_cast<StateLog*>(g)) {
                s->gathered_data.clear();
                s->reorder_index_data.clear();
            }
        } else {
            g->log(step);
        }
    }
}

void IOManager::write_restart_file(long step, std::string dir_name) {
    (void)step;
    if (!state_log) return;
    std::filesystem::path dir = system_dir_path / dir_name;
    make_dir(dir.string(), true);
    orchestrator.reset_dependency_status();
    state_log->handle_dependencies();
    state_log->gather_data(0);
    state_log->write_state_to_path(dir);
    state_log->gathered_data.clear();
}

void IOManager::write_log_configs(std::filesystem::path path) {
    for (auto& c : log_configs)
        c.save(path / (c.at("group_name").get<std::string>() + "_log_config.json"));
}

void IOManager::write_particle_config(std::filesystem::path path) {
    particle.getConfig().save(path / "particle_config.json");
}

void IOManager::write_integrator_config(std::filesystem::path path) {
    if (!integrator) return;
    ConfigDict c = integrator->config;
    c.save(path / "integrator_config.json");
}

void IOManager::write_params() {
    if (system_dir_path.empty()) {
        init_path(system_dir_path, system_dir_name);
        make_dir(system_dir_path.string(), overwrite);
    }
    write_log_configs(system_dir_path);
    write_particle_config(system_dir_path);
    if (integrator) write_integrator_config(system_dir_path);
}

// ---------------------------------------------------------------- load_configs (arg_parsing.h:7-63)
std::tuple<std::unique_ptr<Particle>, long, ConfigDict, ConfigDict, ConfigDict, ConfigDict>
load_configs(int argc, char** argv, bool use_restart) {
    if (argc < 2)
        throw std::invalid_argument(std::string("Usage: ") + (argc > 0 ? argv[0] : "driver") +
                                    " <script_path>");
    std::filesystem::path root = argv[1];
    ConfigDict run_config = load_config_dict(root / "system" / "run_config.json");
    bool minimize = run_config["minimize"];
    std::filesystem::path input_path = run_config["input_dir"].get<std::filesystem::path>();
    std::filesystem::path output_path = run_config["output_dir"].get<std::filesystem::path>();
    std::string load_source = use_restart ? "restart" : "init";
    std::unique_ptr<Particle> particle;
    long step = 0;
    std::filesystem::path config_root;
    if (input_path.empty() && !output_path.empty()) {
        ConfigDict pc = load_config_dict(output_path / "system" / "particle_config.json");
        particle = createParticle(pc, minimize);
        config_root = output_path;
    } else if (!input_path.empty() && !output_path.empty()) {
        config_root = output_path;
        std::tie(particle, step) = loadParticle(input_path, load_source, -2);
        step = 0;
    } else if (!input_path.empty() && output_path.empty()) {
        config_root = input_path;
        std::tie(particle, step) = loadParticle(input_path, load_source, -2);
    } else {
        throw std::invalid_argument(
            "Run config must define either an input or output path!  run_config: " +
            run_config.dump(4));
    }
    ConfigDict console_config, energy_config, state_config;
    console_config.load(config_root / "system" / "console_log_config.json");
    energy_config.load(config_root / "system" / "energy_log_config.json");
    state_config.load(config_root / "system" / "state_log_config.json");
    std::cout << "Done loading data" << std::endl;
    return std::make_tuple(std::move(particle), step, console_config, energy_config, state_config,
                           run_config);
}
