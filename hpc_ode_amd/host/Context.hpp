// Shared execution context and logger: the part of the reference's Context
// (include/ODEIntegrator/Context/Context.hpp:13-66) and BaseLogger
// (include/ODEIntegrator/Logger/BaseLogger.hpp:8-40) that an Integrator uses —
// getMesh(), getLogger(), and info/debug/warning/error on the logger.
// The reference's OutFileService (timestamped output folders) is not part of the
// integration path; a Context here carries an output folder name instead.
#pragma once

#include <iostream>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "Mesh.hpp"

enum class LogLevel { DEBUG = 0, INFO = 1, WARNING = 2, ERROR = 3 };

class BaseLogger {
 public:
  explicit BaseLogger(LogLevel level = LogLevel::INFO) : level_(level) {}
  virtual ~BaseLogger() = default;
  virtual void log(std::string line) = 0;

  void debug(const std::string& m) { emit(LogLevel::DEBUG, "DEBUG", m); }
  void info(const std::string& m) { emit(LogLevel::INFO, "INFO", m); }
  void warning(const std::string& m) { emit(LogLevel::WARNING, "WARNING", m); }
  void error(const std::string& m) { emit(LogLevel::ERROR, "ERROR", m); }

 private:
  void emit(LogLevel l, const char* tag, const std::string& m) {
    if (static_cast<int>(l) < static_cast<int>(level_)) return;
    log(std::string("[") + tag + "] " + m);
  }
  LogLevel level_;
};

class ConsoleLogger : public BaseLogger {
 public:
  explicit ConsoleLogger(LogLevel level = LogLevel::INFO) : BaseLogger(level) {}
  void log(std::string line) override {
    std::lock_guard<std::mutex> g(mu_);
    std::cerr << line << '\n';
  }

 private:
  std::mutex mu_;
};

// Keeps every line in memory (tests read them back).
class MemoryLogger : public BaseLogger {
 public:
  explicit MemoryLogger(LogLevel level = LogLevel::DEBUG) : BaseLogger(level) {}
  void log(std::string line) override {
    std::lock_guard<std::mutex> g(mu_);
    lines.push_back(std::move(line));
  }
  std::vector<std::string> lines;

 private:
  std::mutex mu_;
};

class Context {
 public:
  Context() : logger_(std::make_shared<ConsoleLogger>()), mesh_(std::make_shared<Mesh>()) {}
  explicit Context(std::shared_ptr<Mesh> mesh) : logger_(std::make_shared<ConsoleLogger>()), mesh_(std::move(mesh)) {}
  Context(std::shared_ptr<BaseLogger> logger, std::shared_ptr<Mesh> mesh)
      : logger_(std::move(logger)), mesh_(std::move(mesh)) {}

  void setLogger(std::shared_ptr<BaseLogger> l) { logger_ = std::move(l); }
  void setMesh(std::shared_ptr<Mesh> m) { mesh_ = std::move(m); }
  std::shared_ptr<BaseLogger> getLogger() const { return logger_; }
  std::shared_ptr<Mesh> getMesh() const { return mesh_; }

  std::string outFolder = ".";   // where csvWriter places its file

 private:
  std::shared_ptr<BaseLogger> logger_;
  std::shared_ptr<Mesh> mesh_;
};
