function [dist, idx, time] = gpuknns(plan, queries, data)
% GPUKNNS  execute a plan made by knnsplan.
%   [dist idx time] = gpuknns(plan, queries, data)
%   queries: d x q, data: d x N (one vector per column). Results as gpuknnHeap.
[dist, idx, time] = plan.function(queries, data, plan.k, plan.streams);
end
