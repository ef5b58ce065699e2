"""gRPC ``snn.Snn`` service over libdtb.so (protos/snn.proto:5-219), one
process and one GPU, so the reference client (python/dist_blockwrapper.py)
can drive this engine over the wire unchanged.

protoc is not needed: the messages are declared below as a
FileDescriptorProto with the same package, names, field numbers and types as
protos/snn.proto (the wire format must match), and the service is registered
with grpc's generic handlers.  The handlers map one to one onto the C ABI:

  Init                    load the block files, one block on this GPU
  Run                     dtb_run, one RunResponse per step
  Updateprop              dtb_update_props
  Updatehyperpara         dtb_update_prop_cols (assigned -> assign)
  Updategamma             dtb_gamma_prop_cols
  Updatesample            dtb_set_samples
  Updateoubackgroundparam dtb_set_ou
  GetVersion, Shutdown
  Updatettypecacurrentparam, Updatedopaminecurrentparam,
  Updateadaptationcurrentparam: success = false (not supported here)
"""
import threading
from concurrent import futures

import numpy as np

from . import _lib
from .engine import Block, load_npz

_F = None   # FieldDescriptorProto


def _descriptor():
    from google.protobuf import descriptor_pb2
    global _F
    _F = descriptor_pb2.FieldDescriptorProto
    fd = descriptor_pb2.FileDescriptorProto(name="snn.proto", package="snn", syntax="proto3")
    T = {"int32": _F.TYPE_INT32, "float": _F.TYPE_FLOAT, "double": _F.TYPE_DOUBLE, "bool": _F.TYPE_BOOL,
         "string": _F.TYPE_STRING, "bytes": _F.TYPE_BYTES}

    def enum(name, values, parent=None):
        e = (parent.enum_type if parent is not None else fd.enum_type).add(name=name)
        for k, v in values:
            e.value.add(name=k, number=v)

    def msg(name, fields):
        m = fd.message_type.add(name=name)
        for num, (fname, ftype, rep) in enumerate(fields, start=1):
            f = m.field.add(name=fname, number=num,
                            label=_F.LABEL_REPEATED if rep else _F.LABEL_OPTIONAL)
            if ftype in T:
                f.type = T[ftype]
            elif ftype.startswith("enum:"):
                f.type = _F.TYPE_ENUM
                f.type_name = ftype[5:]
            else:
                f.type = _F.TYPE_MESSAGE
                f.type_name = ftype
        return m

    status = [("SNN_OK", 0), ("SNN_INVALID_PARAMETER", 1), ("SNN_UNSET_SAMPLE_FAILURE", 2),
              ("SNN_UNSUPPORTED", 3), ("SNN_NOT_ENABLE_METRIC", 4), ("SNN_INVALID_NETWORK", 5)]
    enum("SnnStatus", status)
    prop = ["EXT_STIMULI_I", "MEMBRANE_C", "REF_T", "LEAKAGE_G", "LEAKAGE_V", "THRESHOLD_V", "RESET_V",
            "CONDUCT_G_AMPA", "CONDUCT_G_NMDA", "CONDUCT_G_GABAa", "CONDUCT_G_GABAb", "MEMBRANE_V_AMPA",
            "MEMBRANE_V_NMDA", "MEMBRANE_V_GABAa", "MEMBRANE_V_GABAb", "TAO_AMPA", "TAO_NMDA", "TAO_GABAa",
            "TAO_GABAb", "NOISE_RATE"]
    enum("PropType", [(p, i) for i, p in enumerate(prop)])
    msg("Empty", [])
    msg("Version", [("version", "int32", False)])
    init = msg("InitRequest", [("file_path", "string", False), ("delta_t", "float", False),
                               ("mode", "enum:.snn.InitRequest.CommMode", False)])
    enum("CommMode", [("COMM_P2P", 0), ("COMM_ROUTE", 1)], parent=init)
    msg("SubblockInfo", [("subblk_id", "int32", False), ("subblk_num", "int32", False)])
    msg("InitResponse", [("status", "enum:.snn.SnnStatus", False), ("block_id", "int32", False),
                         ("neurons_per_block", "int32", False), ("subblk_info", ".snn.SubblockInfo", True),
                         ("used_cpu_mem", "double", False), ("total_gpu_mem", "double", False),
                         ("used_gpu_mem", "double", False)])
    msg("RunRequest", [("iter", "int32", False), ("iter_offset", "int32", False), ("t_steps", "int32", False),
                       ("output_freq", "bool", False), ("freq_char", "bool", False), ("output_vmean", "bool", False),
                       ("output_imean", "bool", False), ("use_ou_background", "bool", False),
                       ("output_receptor_imeas", "bool", False), ("output_sample_spike", "bool", False),
                       ("output_sample_vmemb", "bool", False), ("output_sample_iou", "bool", False),
                       ("output_sample_isynaptic", "bool", False)])
    msg("RunResponse", [("status", "enum:.snn.SnnStatus", False)] +
        [(n, "bytes", True) for n in ("freq", "vmean", "sample_spike", "sample_vmemb", "sample_iou",
                                      "sample_isynaptic", "imean", "ampa_imean", "nmda_imean", "gabaa_imean",
                                      "gabab_imean")])
    msg("UpdatePropRequest", [("block_id", "int32", False), ("neuron_id", "int32", True),
                              ("prop_id", "enum:.snn.PropType", True), ("prop_val", "float", True)])
    msg("UpdatePropResponse", [("success", "bool", False)])
    msg("UpdateGammaRequest", [("block_id", "int32", False), ("prop_id", "enum:.snn.PropType", True),
                               ("brain_id", "int32", True), ("gamma_concentration", "float", True),
                               ("gamma_rate", "float", True)])
    msg("UpdateGammaResponse", [("success", "bool", False)])
    msg("UpdateHyperParaRequest", [("block_id", "int32", False), ("prop_id", "enum:.snn.PropType", True),
                                   ("brain_id", "int32", True), ("hpara_val", "float", True),
                                   ("assigned", "bool", False)])
    msg("UpdateHyperParaResponse", [("success", "bool", False)])
    msg("UpdateSampleRequest", [("block_id", "int32", False), ("sample_idx", "int32", True)])
    msg("UpdateSampleResponse", [("success", "bool", False)])
    msg("UpdateOUBackgroundParamRequest", [("block_id", "int32", False), ("brain_id", "int32", True),
                                           ("mean", "float", True), ("deviation", "float", True),
                                           ("correlation_time", "float", True)])
    msg("UpdateOUBackgroundParamResponse", [("success", "bool", False)])
    msg("UpdateTTypeCaCurrentParamRequest", [("block_id", "int32", False), ("brain_id", "int32", True)] +
        [(n, "float", True) for n in ("h_init", "g_t", "tao_h_minus", "tao_h_plus", "v_h", "v_t")])
    msg("UpdateTTypeCaCurrentParamResponse", [("success", "bool", False)])
    msg("UpdateDopamineCurrentParamRequest", [("block_id", "int32", False), ("brain_id", "int32", True),
                                              ("v_da", "float", True), ("g_da", "float", True)])
    msg("UpdateDopamineCurrentParamResponse", [("success", "bool", False)])
    msg("UpdateAdaptationCurrentParamRequest", [("block_id", "int32", False), ("brain_id", "int32", True)] +
        [(n, "float", True) for n in ("ca_init", "ca_decay", "alpha_constant", "v_k", "g_ahp")])
    msg("UpdateAdaptationCurrentParamResponse", [("success", "bool", False)])
    msg("ShutdownRequest", [])
    msg("ShutdownResponse", [("shutdown", "bool", False)])
    return fd


class Messages:
    """Message classes of package snn, by name (snn.proto:20-219)."""

    def __init__(self):
        from google.protobuf import descriptor_pool, message_factory
        pool = descriptor_pool.DescriptorPool()
        fd = pool.Add(_descriptor())
        for name in fd.message_types_by_name:
            setattr(self, name, message_factory.GetMessageClass(pool.FindMessageTypeByName("snn." + name)))


# (method, request stream, response stream, request type, response type); snn.proto:5-18
METHODS = [
    ("GetVersion", False, False, "Empty", "Version"),
    ("Init", False, True, "InitRequest", "InitResponse"),
    ("Run", False, True, "RunRequest", "RunResponse"),
    ("Updateprop", True, False, "UpdatePropRequest", "UpdatePropResponse"),
    ("Updategamma", True, False, "UpdateGammaRequest", "UpdateGammaResponse"),
    ("Updatehyperpara", True, False, "UpdateHyperParaRequest", "UpdateHyperParaResponse"),
    ("Updatesample", True, False, "UpdateSampleRequest", "UpdateSampleResponse"),
    ("Updateoubackgroundparam", True, False, "UpdateOUBackgroundParamRequest", "UpdateOUBackgroundParamResponse"),
    ("Updatettypecacurrentparam", True, False, "UpdateTTypeCaCurrentParamRequest",
     "UpdateTTypeCaCurrentParamResponse"),
    ("Updatedopaminecurrentparam", True, False, "UpdateDopamineCurrentParamRequest",
     "UpdateDopamineCurrentParamResponse"),
    ("Updateadaptationcurrentparam", True, False, "UpdateAdaptationCurrentParamRequest",
     "UpdateAdaptationCurrentParamResponse"),
    ("Shutdown", False, False, "ShutdownRequest", "ShutdownResponse"),
]


class SnnServicer:
    """The simulator node of tools/dist_simulator.cpp:3118-3976 for one block
    on one GPU, calling libdtb in-process."""

    def __init__(self, msgs, device=0, seed=0):
        self.m = msgs
        self.device = device
        self.seed = seed
        self.block = None
        self.stop_event = threading.Event()
        # One RPC at a time touches the block (the engine is driven by one
        # thread at a time; dist_simulator.cpp:3993-4047 has one command
        # thread per GPU).
        self._lock = threading.Lock()
        self.run_chunk = 100     # steps per streamed chunk of a Run
        self.run_started = self.run_finished = 0

    def GetVersion(self, request, context):
        return self.m.Version(version=_lib.load().dtb_version())

    def Init(self, request, context):
        with self._lock:
            if self.block is not None:
                self.block.engine.close()
                self.block = None
            try:
                net, _ = load_npz(request.file_path)
                self.block = Block(net, delta_t=request.delta_t or 1.0, seed=self.seed, device=self.device)
            except _lib.DtbError as e:
                status = e.status if e.status <= 5 else 1
                yield self.m.InitResponse(status=status)
                return
            info = [self.m.SubblockInfo(subblk_id=int(p), subblk_num=int(c))
                    for p, c in zip(self.block.subblk_id, self.block.neurons_per_subblk)]
            yield self.m.InitResponse(status=0, block_id=0, neurons_per_block=int(self.block.total_neurons),
                                      subblk_info=info)

    def Run(self, request, context):
        """One RunResponse per step, streamed while the block steps
        (dist_simulator.cpp:3343-3455): the run goes on in chunks of
        run_chunk steps and the responses of a chunk are sent while the next
        one is stepped.  The block is held for the whole run: an Init or
        update arriving meanwhile waits for it."""
        with self._lock:
            b = self.block
            if b is None:
                yield self.m.RunResponse(status=1)
                return
            self.run_started += 1
            stream = b.run_stream(request.iter, chunk=self.run_chunk, freqs=request.output_freq,
                                  output_sample_spike=request.output_sample_spike,
                                  output_sample_vmemb=request.output_sample_vmemb,
                                  output_sample_iou=request.output_sample_iou,
                                  output_sample_isynaptic=request.output_sample_isynaptic,
                                  iou=request.use_ou_background, t_steps=max(1, request.t_steps),
                                  vmean=request.output_vmean, imean=request.output_imean,
                                  receptor_imean=request.output_receptor_imeas)
            try:
                for _, row in stream:
                    r = self.m.RunResponse(status=0)
                    freq = row.get("freq")
                    if freq is not None and request.freq_char:
                        r.freq.append(b.freq_char(freq[None, :])[0].tobytes())   # stat_result.cpp:103-130
                    elif freq is not None:
                        r.freq.append(freq.astype(np.int32).tobytes())
                    for key, field in (("vmean", "vmean"), ("spike", "sample_spike"), ("vmemb", "sample_vmemb"),
                                       ("iou", "sample_iou"), ("isynaptic", "sample_isynaptic"), ("imean", "imean")):
                        if key in row:
                            getattr(r, field).append(row[key].tobytes())
                    rec = row.get("receptor_imean")
                    if rec is not None:
                        for j, field in enumerate(("ampa_imean", "nmda_imean", "gabaa_imean", "gabab_imean")):
                            getattr(r, field).append(rec[j].tobytes())
                    yield r
            except _lib.DtbError as e:
                yield self.m.RunResponse(status=e.status if e.status <= 5 else 1)
            finally:
                stream.close()
                self.run_finished += 1

    def _collect(self, it, fields):
        cols = {f: [] for f in fields}
        for req in it:
            for f in fields:
                cols[f].extend(getattr(req, f))
        return cols

    def Updateprop(self, it, context):
        with self._lock:
            c = self._collect(it, ("neuron_id", "prop_id", "prop_val"))
            ok = self.block is not None
            if ok and c["neuron_id"]:
                try:
                    self.block.engine.update_props(self.block.inst, np.asarray(c["neuron_id"], np.uint32),
                                                   np.asarray(c["prop_id"], np.uint32),
                                                   np.asarray(c["prop_val"], np.float32))
                except _lib.DtbError:
                    ok = False
            return self.m.UpdatePropResponse(success=ok)

    def Updatehyperpara(self, it, context):
        with self._lock:
            ok = self.block is not None
            for req in it:
                if not ok or not req.brain_id:
                    continue
                try:
                    self.block.engine.update_prop_cols(self.block.inst, np.asarray(req.prop_id, np.uint32),
                                                       np.asarray(req.brain_id, np.uint32),
                                                       np.asarray(req.hpara_val, np.float32), assign=req.assigned)
                except _lib.DtbError:
                    ok = False
            return self.m.UpdateHyperParaResponse(success=ok)

    def Updatesample(self, it, context):
        with self._lock:
            c = self._collect(it, ("sample_idx",))
            ok = self.block is not None
            if ok:
                try:
                    self.block.set_samples(np.asarray(c["sample_idx"], np.uint32))
                except _lib.DtbError:
                    ok = False
            return self.m.UpdateSampleResponse(success=ok)

    def Updateoubackgroundparam(self, it, context):
        with self._lock:
            c = self._collect(it, ("brain_id", "mean", "deviation", "correlation_time"))
            ok = self.block is not None
            if ok and c["brain_id"]:
                try:
                    self.block.update_ou_background_stimuli(np.asarray(c["brain_id"], np.uint32),
                                                            np.asarray(c["correlation_time"], np.float32),
                                                            np.asarray(c["mean"], np.float32),
                                                            np.asarray(c["deviation"], np.float32))
                except _lib.DtbError:
                    ok = False
            return self.m.UpdateOUBackgroundParamResponse(success=ok)

    def Updategamma(self, it, context):
        with self._lock:
            ok = self.block is not None
            for req in it:
                if not ok or not req.brain_id:
                    continue
                try:
                    self.block.engine.gamma_prop_cols(self.block.inst, np.asarray(req.prop_id, np.uint32),
                                                      np.asarray(req.brain_id, np.uint32),
                                                      np.asarray(req.gamma_concentration, np.float32),
                                                      np.asarray(req.gamma_rate, np.float32))
                except _lib.DtbError:
                    ok = False
            return self.m.UpdateGammaResponse(success=ok)

    def _unsupported(self, resp):
        def handler(it, context):
            for _ in it:
                pass
            return getattr(self.m, resp)(success=False)
        return handler

    def Shutdown(self, request, context):
        self.stop_event.set()
        return self.m.ShutdownResponse(shutdown=True)

    def close(self):
        with self._lock:
            if self.block is not None:
                self.block.engine.close()
                self.block = None


def make_handler(servicer, msgs):
    import grpc
    handlers = {}
    for name, req_stream, resp_stream, req_t, resp_t in METHODS:
        fn = getattr(servicer, name, None)
        if fn is None:
            fn = servicer._unsupported(resp_t)
        des = getattr(msgs, req_t).FromString
        ser = getattr(msgs, resp_t).SerializeToString
        if req_stream and not resp_stream:
            h = grpc.stream_unary_rpc_method_handler(fn, request_deserializer=des, response_serializer=ser)
        elif resp_stream and not req_stream:
            h = grpc.unary_stream_rpc_method_handler(fn, request_deserializer=des, response_serializer=ser)
        else:
            h = grpc.unary_unary_rpc_method_handler(fn, request_deserializer=des, response_serializer=ser)
        handlers[name] = h
    return grpc.method_handlers_generic_handler("snn.Snn", handlers)


def serve(address="127.0.0.1:50051", device=0, seed=0, max_message_length=1 << 30):
    """Starts the service; returns (server, servicer).  Call
    ``servicer.stop_event.wait()`` then ``server.stop(None)`` to run until a
    Shutdown request."""
    import grpc
    msgs = Messages()
    servicer = SnnServicer(msgs, device=device, seed=seed)
    server = grpc.server(futures.ThreadPoolExecutor(max_workers=4),
                         options=[("grpc.max_send_message_length", max_message_length),
                                  ("grpc.max_receive_message_length", max_message_length)])
    server.add_generic_rpc_handlers((make_handler(servicer, msgs),))
    server.add_insecure_port(address)
    server.start()
    return server, servicer


class SnnClient:
    """Minimal client stub over the same messages (what snn_pb2_grpc.SnnStub
    provides to the reference client)."""

    def __init__(self, channel, msgs=None):
        self.m = msgs or Messages()
        for name, req_stream, resp_stream, req_t, resp_t in METHODS:
            ser = getattr(self.m, req_t).SerializeToString
            des = getattr(self.m, resp_t).FromString
            path = "/snn.Snn/" + name
            if req_stream and not resp_stream:
                call = channel.stream_unary(path, request_serializer=ser, response_deserializer=des)
            elif resp_stream and not req_stream:
                call = channel.unary_stream(path, request_serializer=ser, response_deserializer=des)
            else:
                call = channel.unary_unary(path, request_serializer=ser, response_deserializer=des)
            setattr(self, name, call)


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="snn.Snn gRPC service on one GPU")
    ap.add_argument("--address", default="127.0.0.1:50051")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args()
    srv, svc = serve(a.address, a.device, a.seed)
    svc.stop_event.wait()
    srv.stop(None)
    svc.close()
